"""GPU parity of 3D C2C / Z2Z plans (aclfftPlan3d: one [nx][ny][nz] volume, nz fastest) against
the FP64 oracle's 3D restatement: rows along z, then the compiled column kernels along y and x
(short axes 2 … 32 included), length-1 axes dropped, four-step rows for a long z axis, the
runtime-radix kernel for lengths without a compiled one (complex64), in place and out of place,
both directions.  Tolerances: rel-L2 <= 1e-6 * log2(N) (complex64), 1e-14 * log2(N) (complex128),
N = nx * ny * nz."""
import math

import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

SHAPES_C2C = [(16, 16, 16), (64, 32, 16), (2, 2, 1 << 16), (128, 128, 128), (4, 1, 4096), (1, 1, 1),
              (3, 5, 7), (8, 4, 2), (32, 64, 60), (1, 64, 1)]
SHAPES_Z2Z = [(16, 16, 16), (64, 32, 16), (2, 2, 1 << 16), (128, 64, 128), (4, 1, 4096), (8, 4, 2),
              (32, 64, 60)]


def _run(shape, typ, in_place):
    import torch
    z = typ == aclfft.Z2Z
    x = crandn(shape, sum(shape) + 7 * z)
    if not z:
        x = x.astype(np.complex64)
    p = aclfft.Plan.plan3d(*shape, typ)
    errs = []
    for direction in (aclfft.FORWARD, aclfft.BACKWARD):
        xt = torch.from_numpy(x).cuda()
        yt = xt if in_place else torch.empty_like(xt)
        (p.exec_z2z if z else p.exec_c2c)(xt, yt, direction)
        torch.cuda.synchronize()
        ref = oracle.c2c_3d(x.astype(np.complex128)[None], direction)[0]
        errs.append(oracle.rel_l2(yt.cpu().numpy(), ref))
    print(shape, p.describe(), "rel_l2 %.3e %.3e" % tuple(errs))
    p.destroy()
    return errs


@pytest.mark.parametrize("shape", SHAPES_C2C)
def test_c2c_3d_vs_oracle(cuda, shape):
    n = int(np.prod(shape))
    for err in _run(shape, aclfft.C2C, in_place=False):
        assert err <= fp32_tol(n), err


@pytest.mark.parametrize("shape", SHAPES_Z2Z)
def test_z2z_3d_vs_oracle(cuda, shape):
    n = int(np.prod(shape))
    for err in _run(shape, aclfft.Z2Z, in_place=False):
        assert err <= 1e-14 * max(1.0, math.log2(max(n, 2))), err


@pytest.mark.parametrize("typ", [aclfft.C2C, aclfft.Z2Z])
def test_3d_in_place(cuda, typ):
    shape = (32, 64, 60)
    n = int(np.prod(shape))
    tol = fp32_tol(n) if typ == aclfft.C2C else 1e-14 * math.log2(n)
    for err in _run(shape, typ, in_place=True):
        assert err <= tol, err


def test_3d_round_trip_large(cuda):
    """256^3 complex64 (128 MiB each way): inverse(forward(x)) = N x, checked on the device."""
    import torch
    shape = (256, 256, 256)
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(shape, dtype=torch.complex64, device="cuda", generator=g)
    y = torch.empty_like(x)
    p = aclfft.Plan.plan3d(*shape, aclfft.C2C)
    p.exec_c2c(x, y, aclfft.FORWARD)
    p.exec_c2c(y, y, aclfft.BACKWARD)
    torch.cuda.synchronize()
    n = x.numel()
    err = (torch.linalg.vector_norm(y / n - x) / torch.linalg.vector_norm(x)).item()
    # spot bins against the definition: X[k] for a few k by a direct FP64 sum over the volume
    p.exec_c2c(x, y, aclfft.FORWARD)
    torch.cuda.synchronize()
    xd = x.to(torch.complex128)
    u = torch.arange(256, device="cuda", dtype=torch.float64)
    for k in [(0, 0, 0), (1, 2, 3), (255, 128, 7)]:
        ph = [torch.exp(-2j * math.pi * kk * u / 256) for kk in k]
        ref = ((xd @ ph[2]) @ ph[1] @ ph[0]).item()
        got = complex(y[k].to(torch.complex128).item())
        assert abs(got - ref) <= 1e-4 * math.sqrt(n), (k, got, ref)
    p.destroy()
    assert err <= fp32_tol(n), err
