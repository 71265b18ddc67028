"""GPU: caller-managed workspace (aclfftSetAutoAllocation(0) + aclfftSetWorkArea) - four-step C2C,
large real transforms, in-place real transforms, graph replays across work-area changes, one
area shared by two plans."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu


def _area(torch, nbytes):
    return torch.empty(max(nbytes, 1), dtype=torch.uint8, device="cuda")


def test_fourstep_with_caller_workspace(cuda):
    import torch
    n, b = 1 << 20, 2
    x = crandn((b, n), 21).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.make_plan1d(n, aclfft.C2C, b)
    p.set_auto_allocation(False)
    assert p.get_size() == p.work_size > 0
    with pytest.raises(aclfft.AclfftError) as ei:
        p.exec_c2c(xt, yt, aclfft.FORWARD)
    assert ei.value.code == aclfft.NO_WORKSPACE
    area = _area(torch, p.get_size())
    p.set_work_area(area)
    p.exec_c2c(xt, yt, aclfft.FORWARD)
    torch.cuda.synchronize()
    assert oracle.rel_l2(yt.cpu().numpy(), oracle.c2c(x.astype(np.complex128))) <= fp32_tol(n)
    p.destroy()


@pytest.mark.parametrize("graph", [False, True])
def test_work_area_switch_and_sharing(cuda, graph):
    """Two areas in turn (captured graphs must not keep the old address), then one area shared by
    two plans on one stream."""
    import torch
    n, b = 1 << 18, 3
    x = crandn((b, n), 22).astype(np.complex64)
    ref = oracle.c2c(x.astype(np.complex128), aclfft.BACKWARD)
    xt = torch.from_numpy(x).cuda()
    p = aclfft.Plan.make_plan1d(n, aclfft.C2C, b)
    q = aclfft.Plan.make_plan1d(n, aclfft.C2C, b)
    for pl in (p, q):
        pl.set_auto_allocation(False)
        pl.set_graph_mode(graph)
    a1 = _area(torch, p.get_size())
    a2 = _area(torch, p.get_size())
    for area in (a1, a2, a1):
        a_other = a2 if area is a1 else a1
        a_other.fill_(0xFF)  # stale data in the unused area must not matter
        p.set_work_area(area)
        yt = torch.empty_like(xt)
        p.exec_c2c(xt, yt, aclfft.BACKWARD)
        torch.cuda.synchronize()
        assert oracle.rel_l2(yt.cpu().numpy(), ref) <= fp32_tol(n)
    q.set_work_area(a1)
    y1 = torch.empty_like(xt)
    y2 = torch.empty_like(xt)
    p.exec_c2c(xt, y1, aclfft.BACKWARD)
    q.exec_c2c(xt, y2, aclfft.BACKWARD)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    p.destroy()
    q.destroy()


@pytest.mark.parametrize("n,kind", [(1 << 21, aclfft.R2C), (1 << 21, aclfft.C2R), (4096, aclfft.R2C)])
def test_real_with_caller_workspace_in_place(cuda, n, kind):
    """Large real transforms (four-step on the half-length complex data) and the in-place copy
    of device-memory real transforms both live in the caller's area."""
    import torch
    b = 2
    h = n // 2 + 1
    p = aclfft.Plan.make_plan1d(n, kind, b)
    p.set_auto_allocation(False)
    area = _area(torch, p.get_size())
    p.set_work_area(area)
    buf = torch.zeros(b * 2 * h, dtype=torch.float32, device="cuda")  # in-place: padded real rows
    if kind == aclfft.R2C:
        x = np.random.default_rng(23).standard_normal((b, n)).astype(np.float32)
        bv = buf.view(b, 2 * h)
        bv[:, :n] = torch.from_numpy(x).cuda()
        packed = bv[:, :n].contiguous().view(-1)
        out = torch.empty(b * h, dtype=torch.complex64, device="cuda")
        p.exec_r2c(packed, out)  # out of place on the same plan first
        torch.cuda.synchronize()
        ref = oracle.r2c(x.astype(np.float64))
        assert oracle.rel_l2(out.cpu().numpy().reshape(b, h), ref) <= fp32_tol(n)
        flat = torch.from_numpy(x).cuda().reshape(-1).clone()
        io = torch.zeros(b * h, dtype=torch.complex64, device="cuda")
        io.view(torch.float32)[: b * n] = flat
        p.exec_r2c(io, io)  # in place: input packed rows of n floats, output rows of n/2+1 bins
        torch.cuda.synchronize()
        assert oracle.rel_l2(io.cpu().numpy().reshape(b, h), ref) <= fp32_tol(n)
    else:
        X = crandn((b, h), 24).astype(np.complex64)
        X[:, 0] = X[:, 0].real
        X[:, -1] = X[:, -1].real
        Xt = torch.from_numpy(X).cuda().reshape(-1).clone()
        out = torch.empty(b * n, dtype=torch.float32, device="cuda")
        p.exec_c2r(Xt, out)
        torch.cuda.synchronize()
        ref = oracle.c2r(X.astype(np.complex128), n)
        assert oracle.rel_l2(out.cpu().numpy().reshape(b, n), ref) <= fp32_tol(n)
        p.exec_c2r(Xt, Xt)  # in place
        torch.cuda.synchronize()
        got = Xt.view(torch.float32)[: b * n].cpu().numpy().reshape(b, n)
        assert oracle.rel_l2(got, ref) <= fp32_tol(n)
    p.destroy()
