"""GPU parity: the runtime-radix kernel (k_generic) against the FP64 oracle.

It serves every length without a compiled kernel (any product of the supported radices up to 9216
points) and, with ACLFFT_DISABLE_TEMPLATED=1, every length and layout: contiguous rows, the
transform-major layouts of 2D / 3D columns and VERTICAL plans, real input and Hermitian output.
Batches are odd so that the last workgroup holds fewer transforms than the others.
"""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

GENERIC_LENGTHS = [3, 7, 11, 15, 100, 121, 143, 360, 999, 1200, 1331, 2000, 4913, 9000]


def run_c2c(torch, x, direction):
    b, n = x.shape
    xt = torch.from_numpy(x.astype(np.complex64)).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
    d = p.describe()
    p.exec_c2c(xt, yt, direction)
    torch.cuda.synchronize()
    p.destroy()
    return yt.cpu().numpy(), d


@pytest.mark.parametrize("n", GENERIC_LENGTHS)
@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_generic_lengths(cuda, n, direction):
    import torch
    batch = 97 if n < 1000 else 13
    x = crandn((batch, n), n + direction).astype(np.complex64)
    y, d = run_c2c(torch, x, direction)
    assert "generic" in d, d
    err = oracle.rel_l2(y, oracle.c2c(x.astype(np.complex128), direction))
    print("n=%d %s rel_l2 %.3e" % (n, d, err))
    assert err <= fp32_tol(n)


@pytest.mark.parametrize("n,batch", [(64, 301), (1024, 37), (4096, 5), (6, 1001)])
def test_generic_forced_for_compiled_lengths(cuda, monkeypatch, n, batch):
    import torch
    monkeypatch.setenv("ACLFFT_DISABLE_TEMPLATED", "1")
    x = crandn((batch, n), n).astype(np.complex64)
    y, d = run_c2c(torch, x, aclfft.FORWARD)
    assert "generic" in d, d
    assert oracle.rel_l2(y, oracle.c2c(x.astype(np.complex128))) <= fp32_tol(n)


@pytest.mark.parametrize("shape", [(3, 64, 48), (2, 100, 7), (5, 15, 121)])
def test_generic_2d_columns(cuda, monkeypatch, shape):
    """2D plans on the runtime-radix kernel: the column step walks a transform-major layout."""
    import torch
    monkeypatch.setenv("ACLFFT_DISABLE_TEMPLATED", "1")
    x = crandn(shape, sum(shape)).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan2d(*shape, aclfft.C2C)
    assert "generic" in p.describe()
    p.exec_c2c(xt, yt, aclfft.BACKWARD)
    torch.cuda.synchronize()
    p.destroy()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c_2d(x.astype(np.complex128), aclfft.BACKWARD))
    assert err <= fp32_tol(shape[1] * shape[2])


def test_generic_3d_and_vertical(cuda, monkeypatch):
    import torch
    monkeypatch.setenv("ACLFFT_DISABLE_TEMPLATED", "1")
    x = crandn((12, 10, 14), 3).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    p = aclfft.Plan.plan3d(12, 10, 14, aclfft.C2C)
    p.exec_c2c(xt, xt, aclfft.FORWARD)
    torch.cuda.synchronize()
    p.destroy()
    assert oracle.rel_l2(xt.cpu().numpy(), oracle.c2c_3d(x.astype(np.complex128))) <= fp32_tol(12 * 10 * 14)
    # VERTICAL plans take power-of-two lengths from 256 (the reference's domain)
    v = crandn((512, 33), 4).astype(np.complex64)
    vt = torch.from_numpy(v).cuda()
    p = aclfft.Plan.plan1d(512, aclfft.C2C, 33, aclfft.VERTICAL)
    assert "generic" in p.describe()
    p.exec_c2c(vt, vt, aclfft.FORWARD)
    torch.cuda.synchronize()
    p.destroy()
    assert oracle.rel_l2(vt.cpu().numpy(), oracle.c2c_vertical(v.astype(np.complex128))) <= fp32_tol(512)


@pytest.mark.parametrize("n", [15, 255, 999, 1000, 2002])
def test_generic_real(cuda, monkeypatch, n):
    """R2C (real input, Hermitian half output) and C2R (Hermitian input, real output) on the
    runtime-radix kernel, odd and even lengths."""
    import torch
    monkeypatch.setenv("ACLFFT_DISABLE_TEMPLATED", "1")
    b = 29
    x = np.random.default_rng(n).standard_normal((b, n)).astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    X = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device="cuda")
    back = torch.empty_like(xt)
    pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
    pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
    pr.exec_r2c(xt, X)
    pc.exec_c2r(X, back)
    torch.cuda.synchronize()
    Xh = X.cpu().numpy()
    assert oracle.rel_l2(Xh, oracle.r2c(x.astype(np.float64))) <= fp32_tol(n)
    assert oracle.rel_l2(back.cpu().numpy(), oracle.c2r(Xh.astype(np.complex128), n)) <= fp32_tol(n)
    pr.destroy()
    pc.destroy()
