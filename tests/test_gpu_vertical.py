"""GPU parity of the VERTICAL layout over the reference's whole vertical domain (powers of two
2^8 … 2^18, lib/fft_plan_validate.h:30-47): [n][cols] row-major, one transform per column, in
both precisions, both directions, in place and out of place.  Columns up to 4096 points run in
one column kernel; longer columns as the column four-step (two column steps over compiled
factors, the first with the twiddle W_n^(k1*n2)).  Tolerances: rel-L2 <= 1e-6 * log2(n)
(complex64), 1e-14 * log2(n) (complex128) against the FP64 oracle."""
import math

import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

CASES = [(n, cols) for n in (1 << 8, 1 << 12, 1 << 13, 1 << 14, 1 << 15, 1 << 16) for cols in (3, 40)] + \
        [(1 << 17, 5), (1 << 18, 3)]


@pytest.mark.parametrize("typ", [aclfft.C2C, aclfft.Z2Z])
@pytest.mark.parametrize("n,cols", CASES)
def test_vertical_domain_vs_oracle(cuda, typ, n, cols):
    import torch
    z = typ == aclfft.Z2Z
    v = crandn((n, cols), n + cols + z)
    if not z:
        v = v.astype(np.complex64)
    p = aclfft.Plan.plan1d(n, typ, cols, aclfft.VERTICAL)
    tol = 1e-14 * math.log2(n) if z else fp32_tol(n)
    for direction, in_place in ((aclfft.FORWARD, False), (aclfft.BACKWARD, True)):
        vt = torch.from_numpy(v).cuda()
        out = vt if in_place else torch.empty_like(vt)
        (p.exec_z2z if z else p.exec_c2c)(vt, out, direction)
        torch.cuda.synchronize()
        err = oracle.rel_l2(out.cpu().numpy(), oracle.c2c_vertical(v.astype(np.complex128), direction))
        print(p.describe(), "dir", direction, "rel_l2 %.3e" % err)
        assert err <= tol, err
    p.destroy()
