"""GPU parity: aclfftExecC2C (device and host pointers) against the FP64 oracle."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

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
