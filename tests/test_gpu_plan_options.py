"""GPU: plan options across plan forms - the output scale (aclfftExtSetScale) on uneven plans in both
forms, adjoint executions across work-area changes, and device pointers that are not 16-byte
aligned (torch views at an odd complex64 offset, uneven groups at odd offsets)."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu


def check_groups(out, x, groups, direction, scale=1.0):
    off = 0
    for n, b in groups:
        ref = oracle.c2c(x[off:off + n * b].reshape(b, n).astype(np.complex128), direction) * scale
        err = oracle.rel_l2(out[off:off + n * b].reshape(b, n), ref)
        assert err <= fp32_tol(n), (n, b, err)
        off += n * b


# lengths with a persistent class, with a tile kernel only, with neither (runtime radix, four-step)
# and length 1
GROUPS = [(60, 3), (420, 5), (1024, 7), (46200, 1), (1, 4), (6561, 2), (1500, 3), (2048, 2)]


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("graph", [False, True])
def test_set_scale_on_uneven_plans(cuda, monkeypatch, mode, graph):
    import torch
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", str(mode))
    ln = [g[0] for g in GROUPS]
    bt = [g[1] for g in GROUPS]
    total = sum(n * b for n, b in GROUPS)
    x = crandn(total, 31).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    p.set_graph_mode(graph)
    p.set_scale(0.125)
    print(p.describe(), flush=True)
    for direction in (aclfft.FORWARD, aclfft.BACKWARD):
        yt.zero_()
        p.exec_c2c(xt, yt, direction)
        torch.cuda.synchronize()
        check_groups(yt.cpu().numpy(), x, GROUPS, direction, 0.125)
    p.destroy()


def test_set_scale_length_one(cuda):
    import torch
    x = crandn((5, 1), 32).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan1d(1, aclfft.C2C, 5)
    p.set_scale(3.0)
    p.exec_c2c(xt, yt, aclfft.FORWARD)
    torch.cuda.synchronize()
    assert np.allclose(yt.cpu().numpy(), 3.0 * x, rtol=1e-6, atol=0)
    p.destroy()


@pytest.mark.parametrize("n", [4096, 1 << 16])
def test_adjoint_after_work_area_change(cuda, n):
    """Adjoint execution, then a new caller work area with the old one freed, then the adjoint
    again: the adjoint must follow the plan's current workspace."""
    import torch
    b = 3
    rng = np.random.default_rng(33)
    g = (rng.standard_normal((b, n // 2 + 1)) + 1j * rng.standard_normal((b, n // 2 + 1))).astype(np.complex64)
    g[:, 0] = g[:, 0].real
    g[:, -1] = g[:, -1].real
    # adjoint of R2C: grad_x[m] = Re sum_k w_k g[k] e^{+2 pi i km/n}, w = 1 at DC / Nyquist, 1/2 between
    w = np.full(n // 2 + 1, 0.5)
    w[0] = w[-1] = 1.0
    ref = oracle.c2r((g * w).astype(np.complex128), n)
    gt = torch.from_numpy(g).cuda()
    out = torch.empty(b, n, device="cuda")
    p = aclfft.Plan.make_plan1d(n, aclfft.R2C, b)
    p.set_auto_allocation(False)
    area1 = torch.empty(p.get_size(), dtype=torch.uint8, device="cuda")
    p.set_work_area(area1)
    p.exec_r2c_adjoint(gt, out)
    torch.cuda.synchronize()
    assert oracle.rel_l2(out.cpu().numpy(), ref) <= fp32_tol(n)
    area2 = torch.empty_like(area1)
    p.set_work_area(area2)
    del area1
    torch.cuda.empty_cache()
    filler = torch.full((1 << 22,), 7.0, device="cuda")  # reuses freed memory, if any
    out.zero_()
    p.exec_r2c_adjoint(gt, out)
    torch.cuda.synchronize()
    assert oracle.rel_l2(out.cpu().numpy(), ref) <= fp32_tol(n)
    del filler
    p.destroy()


@pytest.mark.parametrize("n", [1024, 1000, 60, 4096, 1 << 16])
def test_c2c_on_odd_offset_views(cuda, n):
    """Device pointers 8 bytes past a 16-byte boundary: staged kernels fall back to a direct
    candidate, or the call reports ACLFFT_UNALIGNED_DATA; either way no wrong result."""
    import torch
    b = 6
    x = crandn((b, n), 34).astype(np.complex64)
    xbuf = torch.zeros(b * n + 1, dtype=torch.complex64, device="cuda")
    ybuf = torch.zeros(b * n + 1, dtype=torch.complex64, device="cuda")
    xv = xbuf[1:]
    yv = ybuf[1:]
    xv.copy_(torch.from_numpy(x.reshape(-1)).cuda())
    assert xv.data_ptr() % 16 == 8
    p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
    try:
        p.exec_c2c(xv, yv, aclfft.FORWARD)
    except aclfft.AclfftError as e:
        assert e.code == aclfft.UNALIGNED_DATA
        p.destroy()
        return
    torch.cuda.synchronize()
    assert oracle.rel_l2(yv.cpu().numpy().reshape(b, n), oracle.c2c(x.astype(np.complex128))) <= fp32_tol(n)
    assert ybuf[0].item() == 0  # nothing written before the view
    p.destroy()


@pytest.mark.parametrize("n", [4096, 3000, 8192, 1 << 15])
def test_r2c_c2r_on_odd_offset_views(cuda, n):
    import torch
    b = 4
    rng = np.random.default_rng(35)
    x = rng.standard_normal((b, n)).astype(np.float32)
    h = n // 2 + 1
    xbuf = torch.zeros(b * n + 1, device="cuda")
    Xbuf = torch.zeros(b * h + 1, dtype=torch.complex64, device="cuda")
    xv = xbuf[1:]
    Xv = Xbuf[1:]
    xv.copy_(torch.from_numpy(x.reshape(-1)).cuda())
    pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
    pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
    try:
        pr.exec_r2c(xv, Xv)
    except aclfft.AclfftError as e:
        assert e.code == aclfft.UNALIGNED_DATA
    else:
        torch.cuda.synchronize()
        X = Xv.cpu().numpy().reshape(b, h)
        assert oracle.rel_l2(X, oracle.r2c(x.astype(np.float64))) <= fp32_tol(n)
    X = oracle.r2c(x.astype(np.float64)).astype(np.complex64)
    Xv.copy_(torch.from_numpy(X.reshape(-1)).cuda())
    ybuf = torch.zeros(b * n + 1, device="cuda")
    yv = ybuf[1:]
    try:
        pc.exec_c2r(Xv, yv)
    except aclfft.AclfftError as e:
        assert e.code == aclfft.UNALIGNED_DATA
    else:
        torch.cuda.synchronize()
        assert oracle.rel_l2(yv.cpu().numpy().reshape(b, n), oracle.c2r(X.astype(np.complex128), n)) <= fp32_tol(n)
    pr.destroy()
    pc.destroy()


@pytest.mark.parametrize("mode", [1, 2])
def test_uneven_groups_at_odd_offsets(cuda, monkeypatch, mode):
    """Group offsets that are odd multiples of 8 bytes: (243, 1) puts the 1500-point group 8 bytes
    past a 16-byte boundary."""
    import torch
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", str(mode))
    groups = [(243, 1), (1500, 3), (1024, 3), (60, 1), (4096, 2)]
    ln = [g[0] for g in groups]
    bt = [g[1] for g in groups]
    total = sum(n * b for n, b in groups)
    x = crandn(total, 36).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    print(p.describe(), flush=True)
    try:
        p.exec_c2c(xt, yt, aclfft.FORWARD)
    except aclfft.AclfftError as e:
        assert e.code == aclfft.UNALIGNED_DATA
        p.destroy()
        return
    torch.cuda.synchronize()
    check_groups(yt.cpu().numpy(), x, groups, aclfft.FORWARD)
    p.destroy()


# (type, n, batch): single-launch plans whose output is at least 16 MiB take the streaming-store
# twin of their default kernel
STORE_CASES = [(aclfft.C2C, 4096, 1024), (aclfft.C2C, 1024, 4096), (aclfft.R2C, 8192, 1024),
               (aclfft.C2R, 8192, 1024), (aclfft.R2C, 4096, 2048), (aclfft.C2R, 3000, 2048)]


@pytest.mark.parametrize("kind,n,batch", STORE_CASES)
def test_streaming_store_rule(cuda, monkeypatch, kind, n, batch):
    """Plans of one launch with >= 16 MiB of output write it with streaming stores (the kernel's
    description names "ntstore"); ACLFFT_STREAM_STORES=0 keeps cached stores.  Both against the
    FP64 oracle, and the two outputs bit-identical (only the store instruction differs)."""
    import torch
    outs = []
    for rule in ("1", "0"):
        monkeypatch.setenv("ACLFFT_STREAM_STORES", rule)
        p = aclfft.Plan.plan1d(n, kind, batch)
        d = p.describe()
        assert ("ntstore" in d) == (rule == "1"), d
        if kind == aclfft.C2C:
            x = crandn((batch, n), 3).astype(np.complex64)
            xt = torch.from_numpy(x).cuda()
            yt = torch.empty_like(xt)
            p.exec_c2c(xt, yt, aclfft.FORWARD)
            ref = oracle.c2c(x[:64].astype(np.complex128))
        elif kind == aclfft.R2C:
            x = np.random.default_rng(3).standard_normal((batch, n)).astype(np.float32)
            xt = torch.from_numpy(x).cuda()
            yt = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device=cuda)
            p.exec_r2c(xt, yt)
            ref = oracle.r2c(x[:64].astype(np.float64))
        else:
            X = crandn((batch, n // 2 + 1), 3).astype(np.complex64)
            xt = torch.from_numpy(X).cuda()
            yt = torch.empty((batch, n), dtype=torch.float32, device=cuda)
            p.exec_c2r(xt, yt)
            ref = oracle.c2r(X[:64].astype(np.complex128), n)
        torch.cuda.synchronize()
        y = yt.cpu().numpy()
        assert oracle.rel_l2(y[:64], ref) <= fp32_tol(n)
        outs.append(y)
        p.destroy()
    assert np.array_equal(outs[0], outs[1])
