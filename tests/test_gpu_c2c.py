"""GPU parity: aclfftExecC2C (device and host pointers) against the FP64 oracle."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import col_variants, crandn, fp32_tol

pytestmark = pytest.mark.gpu

C2C_SIZES = sorted({k["n"] for k in aclfft.kernel_registry() if k["kind"] == 0})


def run_c2c_device(torch, n, batch, x, direction):
    xt = torch.from_numpy(x.astype(np.complex64)).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    p.exec_c2c(xt, yt, direction)
    torch.cuda.synchronize()
    out = yt.cpu().numpy()
    p.destroy()
    return out


@pytest.mark.parametrize("n", C2C_SIZES)
@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_c2c_registry_sizes(cuda, n, direction):
    import torch
    batch = 37
    x = crandn((batch, n), seed=n)
    out = run_c2c_device(torch, n, batch, x, direction)
    ref = oracle.c2c(x.astype(np.complex64).astype(np.complex128), direction)
    err = oracle.rel_l2(out, ref)
    print("n=%d dir=%d rel_l2=%.3e" % (n, direction, err))
    assert err <= fp32_tol(n), err


def test_c2c_host_pointers(cuda):
    n, batch = 1024, 64
    x = crandn((batch, n), seed=0).astype(np.complex64)
    y = np.zeros_like(x)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    p.exec_c2c(x, y, aclfft.FORWARD)
    ref = oracle.c2c(x.astype(np.complex128))
    assert oracle.rel_l2(y, ref) <= fp32_tol(n)
    # in place on host memory
    z = x.copy()
    p.exec_c2c(z, z, aclfft.FORWARD)
    assert np.array_equal(z, y)


def test_c2c_in_place_device(cuda):
    import torch
    n, batch = 1024, 300
    x = crandn((batch, n), seed=5).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    p.exec_c2c(xt, xt, aclfft.FORWARD)
    torch.cuda.synchronize()
    ref = oracle.c2c(x.astype(np.complex128))
    assert oracle.rel_l2(xt.cpu().numpy(), ref) <= fp32_tol(n)


def _variants():
    """(kind, n, index among that (kind, n)'s candidates) for every compiled C2C, R2C and C2R
    kernel (the four-step row kernels and the tile-list kernels have tests of their own)."""
    seen = {}
    out = []
    for k in aclfft.kernel_registry():
        if k["kind"] not in (0, 1, 2):
            continue
        key = (k["kind"], k["n"])
        idx = seen.get(key, 0)
        seen[key] = idx + 1
        out.append((k["kind"], k["n"], idx))
    return out


@pytest.mark.parametrize("kind,n,variant", _variants())
def test_every_registry_kernel(cuda, monkeypatch, kind, n, variant):
    """Each compiled configuration (radix plan, threads per transform, staged / paired I/O),
    forced with ACLFFT_KERNEL_VARIANT, against the FP64 oracle; odd batch exercises partial
    workgroups."""
    import torch
    monkeypatch.setenv("ACLFFT_KERNEL_VARIANT", str(variant))
    batch = 21
    if kind == 0:
        x = crandn((batch, n), n + variant).astype(np.complex64)
        p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
        assert "id=" in p.describe()
        xt = torch.from_numpy(x).cuda()
        yt = torch.empty_like(xt)
        p.exec_c2c(xt, yt, aclfft.BACKWARD)
        torch.cuda.synchronize()
        ref = oracle.c2c(x.astype(np.complex128), aclfft.BACKWARD)
        err = oracle.rel_l2(yt.cpu().numpy(), ref)
    elif kind == 1:
        m = 2 * n
        x = np.random.default_rng(variant).standard_normal((batch, m)).astype(np.float32)
        p = aclfft.Plan.plan1d(m, aclfft.R2C, batch)
        xt = torch.from_numpy(x).cuda()
        yt = torch.empty((batch, m // 2 + 1), dtype=torch.complex64, device=cuda)
        p.exec_r2c(xt, yt)
        torch.cuda.synchronize()
        err = oracle.rel_l2(yt.cpu().numpy(), oracle.r2c(x.astype(np.float64)))
    else:
        m = 2 * n
        X = crandn((batch, m // 2 + 1), variant).astype(np.complex64)
        p = aclfft.Plan.plan1d(m, aclfft.C2R, batch)
        Xt = torch.from_numpy(X).cuda()
        yt = torch.empty((batch, m), dtype=torch.float32, device=cuda)
        p.exec_c2r(Xt, yt)
        torch.cuda.synchronize()
        err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2r(X.astype(np.complex128), m))
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    assert err <= fp32_tol(2 * n)


@pytest.mark.parametrize("rows,variant", col_variants(3))
@pytest.mark.parametrize("cols", [40, 7])
def test_every_column_kernel_2d_and_vertical(cuda, monkeypatch, rows, variant, cols):
    """Column kernels (2D second pass and the VERTICAL layout) against the FP64 oracle, with a
    column count that leaves a partial column group."""
    import torch
    monkeypatch.setenv("ACLFFT_COLUMN_VARIANT", str(variant))
    x = crandn((3, rows, cols), rows + cols + variant).astype(np.complex64)
    p = aclfft.Plan.plan2d(3, rows, cols, aclfft.C2C)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p.exec_c2c(xt, yt, aclfft.BACKWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c_2d(x.astype(np.complex128), aclfft.BACKWARD))
    print(p.describe(), "rel_l2 %.3e" % err)
    assert err <= fp32_tol(rows * cols)
    p.destroy()
    if rows >= 256:  # VERTICAL plans exist for power-of-two rows in [256, 2^18]
        v = crandn((rows, cols), variant).astype(np.complex64)
        p = aclfft.Plan.plan1d(rows, aclfft.C2C, cols, aclfft.VERTICAL)
        vt = torch.from_numpy(v).cuda()
        p.exec_c2c(vt, vt, aclfft.FORWARD)
        torch.cuda.synchronize()
        err = oracle.rel_l2(vt.cpu().numpy(), oracle.c2c_vertical(v.astype(np.complex128)))
        assert err <= fp32_tol(rows)
        p.destroy()


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, (1 << 24) + 16 * 37])
def test_stream_copy(cuda, nbytes):
    """ofx_stream_copy (the benchmarks' bandwidth yardstick) copies exactly, including a last
    partial workgroup, and leaves the bytes after the range untouched."""
    import torch
    from aclfft._lib import stream_copy
    src = torch.randint(0, 255, (nbytes + 64,), dtype=torch.uint8, device="cuda")
    dst = torch.zeros_like(src)
    stream_copy(src.data_ptr(), dst.data_ptr(), nbytes)
    torch.cuda.synchronize()
    assert torch.equal(dst[:nbytes], src[:nbytes])
    assert int(dst[nbytes:].abs().sum()) == 0
    with pytest.raises(RuntimeError):
        stream_copy(src.data_ptr(), dst.data_ptr(), 24)  # not a multiple of 16
