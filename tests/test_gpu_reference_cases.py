"""GPU parity on the reference's own case lists (tests/ref_cases.py).

For every case the output of the aclfft C API on the GPU must pass the reference's verdict
(src/common/host/test_helper.py:41-71, the reference's atol/rtol table) against the reference's
golden path (torch.fft on the CPU, same seed and distribution as its gen_data.py), and, where the
FP64 oracle is affordable, meet this build's own bar rel-L2 <= 1e-6*log2(n) against the oracle.
"""
import math

import numpy as np
import pytest

import aclfft
from oracle import oracle
import ref_cases as rc
from _util import fp32_tol

pytestmark = pytest.mark.gpu

ORACLE_MAX_ELEMS = 1 << 22


def _verdict(name, out, golden, atol, rtol):
    ok, detail = oracle.comparison_details(out, golden, atol, rtol)
    print("%s: %s" % (name, detail))
    assert ok, "%s: %s" % (name, detail)


@pytest.mark.parametrize("case", rc.C2C, ids=[c[0] for c in rc.C2C])
def test_reference_c2c(cuda, case):
    import torch
    name, n, batch, mode, atol, rtol = case
    torch.manual_seed(rc.SEED)
    x = torch.randn(batch, n, dtype=torch.complex64)
    xd = x.to(cuda)
    yd = torch.empty_like(xd)
    if mode == "vertical":
        plan = aclfft.Plan.plan1d(n, aclfft.C2C, batch, aclfft.VERTICAL)
        plan.exec_c2c(xd, yd, aclfft.FORWARD)
        golden = torch.fft.fft(x.reshape(n, batch), dim=0).reshape(batch, n)
        ref_fn = lambda a: oracle.c2c_vertical(a.reshape(n, batch)).reshape(batch, n)
    elif mode == "roundtrip":
        plan = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
        plan.exec_c2c(xd, yd, aclfft.FORWARD)
        plan.exec_c2c(yd, yd, aclfft.BACKWARD)
        golden = torch.fft.ifft(torch.fft.fft(x, dim=1), dim=1) * n
        ref_fn = lambda a: a * n
    else:
        direction = aclfft.FORWARD if mode == "forward" else aclfft.BACKWARD
        plan = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
        plan.exec_c2c(xd, yd, direction)
        golden = torch.fft.fft(x, dim=1) if mode == "forward" else torch.fft.ifft(x, dim=1) * n
        ref_fn = lambda a: oracle.c2c(a, direction)
    torch.cuda.synchronize()
    out = yd.cpu().numpy()
    _verdict(name, out, golden.numpy(), atol, rtol)
    if n * batch <= ORACLE_MAX_ELEMS:
        ref = ref_fn(x.numpy().astype(np.complex128))
        err = oracle.rel_l2(out, ref)
        print("%s: rel_l2 vs FP64 oracle %.3e" % (name, err))
        assert err <= fp32_tol(n) * (2 if mode == "roundtrip" else 1)
    plan.destroy()


@pytest.mark.parametrize("case", rc.R2C, ids=[c[0] for c in rc.R2C])
def test_reference_r2c(cuda, case):
    import torch
    name, n, batch, atol, rtol = case
    torch.manual_seed(rc.SEED)
    x = torch.randn(batch * n).float().reshape(batch, n)
    xd = x.to(cuda)
    yd = torch.empty(batch, n // 2 + 1, dtype=torch.complex64, device=cuda)
    plan = aclfft.Plan.plan1d(n, aclfft.R2C, batch)
    plan.exec_r2c(xd, yd)
    torch.cuda.synchronize()
    del xd
    out = yd.cpu().numpy()
    del yd
    golden = torch.fft.rfft(x, dim=1).numpy()
    _verdict(name, out, golden, atol, rtol)
    # Nyquist / DC imaginary parts are exact zeros (reference rfft1_d_test_arch32.cpp:45-81)
    assert np.all(out[:, 0].imag == 0)
    if n % 2 == 0:
        assert np.all(out[:, -1].imag == 0)
    if n * batch <= ORACLE_MAX_ELEMS:
        err = oracle.rel_l2(out, oracle.r2c(x.numpy().astype(np.float64)))
        print("%s: rel_l2 vs FP64 oracle %.3e" % (name, err))
        assert err <= fp32_tol(n)
    plan.destroy()


@pytest.mark.parametrize("case", rc.C2R, ids=[c[0] for c in rc.C2R])
def test_reference_c2r(cuda, case):
    import torch
    name, n, batch, nyquist_only, atol, rtol = case
    h = n // 2 + 1
    torch.manual_seed(rc.SEED)
    if nyquist_only:
        X = torch.zeros(batch, h, dtype=torch.complex64)
        X[:, -1] = 1.0
    else:
        X = torch.randn(batch, h, dtype=torch.complex64)
    Xd = X.to(cuda)
    yd = torch.empty(batch, n, dtype=torch.float32, device=cuda)
    plan = aclfft.Plan.plan1d(n, aclfft.C2R, batch)
    plan.exec_c2r(Xd, yd)
    torch.cuda.synchronize()
    out = yd.cpu().numpy()
    golden = (torch.fft.irfft(X, n=n, dim=1) * n).numpy()
    _verdict(name, out, golden, atol, rtol)
    if nyquist_only:  # closed form: x[m] = (-1)^m
        assert np.allclose(out, np.where(np.arange(n) % 2 == 0, 1.0, -1.0)[None, :], atol=1e-5)
    if n * batch <= ORACLE_MAX_ELEMS:
        err = oracle.rel_l2(out, oracle.c2r(X.numpy().astype(np.complex128), n))
        print("%s: rel_l2 vs FP64 oracle %.3e" % (name, err))
        assert err <= fp32_tol(n)
    plan.destroy()


@pytest.mark.parametrize("case", rc.FFT2, ids=[c[0] for c in rc.FFT2])
def test_reference_fft2(cuda, case):
    import torch
    name, nx, ny, batch, forward, atol, rtol = case
    torch.manual_seed(rc.SEED)
    x = torch.randn(batch, nx, ny, dtype=torch.complex64)
    xd = x.to(cuda)
    yd = torch.empty_like(xd)
    plan = aclfft.Plan.plan2d(batch, nx, ny, aclfft.C2C)
    direction = aclfft.FORWARD if forward else aclfft.BACKWARD
    plan.exec_c2c(xd, yd, direction)
    torch.cuda.synchronize()
    out = yd.cpu().numpy()
    golden = torch.fft.fft2(x, dim=(1, 2)) if forward else torch.fft.ifft2(x, dim=(1, 2)) * (nx * ny)
    _verdict(name, out, golden.numpy(), atol, rtol)
    err = oracle.rel_l2(out, oracle.c2c_2d(x.numpy().astype(np.complex128), direction))
    print("%s: rel_l2 vs FP64 oracle %.3e" % (name, err))
    assert err <= fp32_tol(nx * ny)
    plan.destroy()
