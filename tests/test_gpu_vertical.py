"""GPU parity of long column transforms: the VERTICAL layout over the reference's whole vertical
domain, 2D plans with long columns and a 3D plan with a long slowest axis.  VERTICAL: (powers of two
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


@pytest.mark.parametrize("typ", [aclfft.C2C, aclfft.Z2Z])
@pytest.mark.parametrize("shape", [(2, 16384, 64), (2, 1 << 17, 8), (1, 8192, 1024), (3, 8192, 5)])
def test_long_columns_2d_vs_oracle(cuda, typ, shape):
    """2D plans whose column length has no single column kernel: the column four-step over
    several images (the second step's output image index split into (image, k1))."""
    import torch
    z = typ == aclfft.Z2Z
    b, nx, ny = shape
    x = crandn(shape, nx + ny + b)
    if not z:
        x = x.astype(np.complex64)
    p = aclfft.Plan.plan2d(b, nx, ny, typ)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    (p.exec_z2z if z else p.exec_c2c)(xt, yt, aclfft.FORWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c_2d(x.astype(np.complex128)))
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    tol = 1e-14 * math.log2(nx * ny) if z else fp32_tol(nx * ny)
    assert err <= tol, err


def test_long_x_axis_3d_vs_oracle(cuda):
    import torch
    shape = (16384, 4, 8)
    x = crandn(shape, 77).astype(np.complex64)
    p = aclfft.Plan.plan3d(*shape, aclfft.C2C)
    xt = torch.from_numpy(x).cuda()
    p.exec_c2c(xt, xt, aclfft.BACKWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(xt.cpu().numpy(), oracle.c2c_3d(x.astype(np.complex128)[None], aclfft.BACKWARD)[0])
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    assert err <= fp32_tol(int(np.prod(shape))), err
