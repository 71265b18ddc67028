"""GPU parity: aclfftExecR2C / aclfftExecC2R against the FP64 oracle."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

# real lengths n whose half (even n) has a compiled kernel
REAL_SIZES = sorted({2 * k["n"] for k in aclfft.kernel_registry() if k["kind"] == 1})


@pytest.mark.parametrize("n", REAL_SIZES)
def test_r2c(cuda, n):
    import torch
    batch = 19
    rng = np.random.default_rng(n)
    x = rng.standard_normal((batch, n)).astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device=cuda)
    p = aclfft.Plan.plan1d(n, aclfft.R2C, batch)
    p.exec_r2c(xt, yt)
    torch.cuda.synchronize()
    out = yt.cpu().numpy()
    ref = oracle.r2c(x.astype(np.float64))
    err = oracle.rel_l2(out, ref)
    print("r2c n=%d rel_l2=%.3e" % (n, err))
    assert err <= fp32_tol(n)
    # DC and Nyquist imaginary parts are exact zeros (reference rfft1_d_test_arch32.cpp:45-81)
    assert np.all(out[:, 0].imag == 0) and np.all(out[:, -1].imag == 0)


@pytest.mark.parametrize("n", REAL_SIZES)
def test_c2r(cuda, n):
    import torch
    batch = 19
    X = crandn((batch, n // 2 + 1), seed=n + 1).astype(np.complex64)
    Xt = torch.from_numpy(X).cuda()
    yt = torch.empty((batch, n), dtype=torch.float32, device=cuda)
    p = aclfft.Plan.plan1d(n, aclfft.C2R, batch)
    p.exec_c2r(Xt, yt)
    torch.cuda.synchronize()
    out = yt.cpu().numpy()
    ref = oracle.c2r(X.astype(np.complex128), n)
    err = oracle.rel_l2(out, ref)
    print("c2r n=%d rel_l2=%.3e" % (n, err))
    assert err <= fp32_tol(n)
    # input untouched (out-of-place)
    assert np.array_equal(Xt.cpu().numpy(), X)
