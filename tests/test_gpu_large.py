"""GPU parity of the large-length C2C paths (four-step: column kernel with fused W_N twiddles, then
row kernel with transposed store; or the runtime-radix kernel for lengths without compiled
factors) against the FP64 oracle, both directions, batch > 1 so that images do not overlap."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

LARGE = [
    (1 << 14, 3), (1 << 15, 3), (1 << 16, 2), (1 << 17, 2), (1 << 18, 2), (1 << 20, 2),
    (1 << 22, 1), (1 << 24, 1), (3 << 14, 2), (1000000, 1), (46200, 3),
]


@pytest.mark.parametrize("n,batch", LARGE)
@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_large_c2c(cuda, n, batch, direction):
    import torch
    x = crandn((batch, n), seed=n % 1000003 + direction).astype(np.complex64)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p.exec_c2c(xt, yt, direction)
    torch.cuda.synchronize()
    desc = p.describe()
    p.destroy()
    ref = oracle.c2c(x.astype(np.complex128), direction)
    err = oracle.rel_l2(yt.cpu().numpy(), ref)
    print("n=%d batch=%d dir=%d rel_l2=%.3e\n%s" % (n, batch, direction, err, desc))
    assert err <= fp32_tol(n), err


@pytest.mark.parametrize("n", [1 << 16, 1 << 20])
def test_large_c2c_in_place(cuda, n):
    """The four-step path with out == in (the row step writes the caller's buffer from the
    plan's workspace)."""
    import torch
    batch = 2
    x = crandn((batch, n), seed=7).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    p.exec_c2c(xt, xt, aclfft.FORWARD)
    torch.cuda.synchronize()
    p.destroy()
    err = oracle.rel_l2(xt.cpu().numpy(), oracle.c2c(x.astype(np.complex128)))
    assert err <= fp32_tol(n), err
