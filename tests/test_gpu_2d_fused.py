"""GPU: the one-pass 2D C2C kernels (images that fit in LDS, nx, ny in {32, 64, 128}) against the
FP64 oracle and against the two-pass form (row kernel + column kernel, ACLFFT_2D_FUSED=0)."""
import ctypes

import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

SIZES = [(nx, ny) for nx in (32, 64, 128) for ny in (32, 64, 128)]


def _fused_sizes():
    lib = aclfft.lib
    lib.ofx_2d_kernel_count.restype = ctypes.c_int
    return lib.ofx_2d_kernel_count()


def test_every_size_has_a_kernel():
    assert _fused_sizes() == len(SIZES)


@pytest.mark.parametrize("nx,ny", SIZES)
@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_fused_2d_vs_oracle(cuda, nx, ny, direction):
    """Odd batch, both directions; the plan must run one fused launch."""
    import torch
    batch = 5
    x = crandn((batch, nx, ny), nx * 3 + ny + (direction > 0)).astype(np.complex64)
    p = aclfft.Plan.plan2d(batch, nx, ny, aclfft.C2C)
    d = p.describe()
    assert d.startswith("fused2d %dx%d" % (nx, ny)) and "|" not in d, d
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p.exec_c2c(xt, yt, direction)
    torch.cuda.synchronize()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c_2d(x.astype(np.complex128), direction))
    print(d, "rel_l2 %.3e" % err)
    assert err <= fp32_tol(nx * ny)
    p.destroy()


@pytest.mark.parametrize("nx,ny", [(64, 64), (32, 128), (128, 32), (128, 128)])
def test_fused_2d_matches_two_pass(cuda, monkeypatch, nx, ny):
    """The fused kernel and the row + column kernels compute the same transform (to FP32
    rounding), in place and out of place."""
    import torch
    batch = 7
    x = crandn((batch, nx, ny), nx + ny).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    p = aclfft.Plan.plan2d(batch, nx, ny, aclfft.C2C)
    y1 = torch.empty_like(xt)
    p.exec_c2c(xt, y1, aclfft.FORWARD)
    z = xt.clone()
    p.exec_c2c(z, z, aclfft.FORWARD)  # in place
    torch.cuda.synchronize()
    p.destroy()
    monkeypatch.setenv("ACLFFT_2D_FUSED", "0")
    q = aclfft.Plan.plan2d(batch, nx, ny, aclfft.C2C)
    assert "fused2d" not in q.describe()
    y2 = torch.empty_like(xt)
    q.exec_c2c(xt, y2, aclfft.FORWARD)
    torch.cuda.synchronize()
    q.destroy()
    assert torch.equal(y1, z)
    a, b = y1.cpu().numpy(), y2.cpu().numpy()
    assert oracle.rel_l2(a, b) <= fp32_tol(nx * ny)


def test_fused_2d_unaligned_scaled_and_3d(cuda):
    """Images that start off a 16-byte boundary (a one-element offset view), the fused output
    scale, and a 3D plan with a unit axis (a 1 x 64 x 64 volume is one 64 x 64 image)."""
    import torch
    nx, ny, batch = 64, 64, 3
    x = crandn((batch * nx * ny + 1,), 9).astype(np.complex64)
    base = torch.from_numpy(x).cuda()
    view = base[1:]
    assert view.data_ptr() % 16 == 8
    y = torch.empty(batch * nx * ny, dtype=torch.complex64, device=cuda)
    p = aclfft.Plan.plan2d(batch, nx, ny, aclfft.C2C)
    p.set_scale(1.0 / (nx * ny))
    p.exec_c2c(view, y, aclfft.BACKWARD)
    torch.cuda.synchronize()
    ref = oracle.c2c_2d(x[1:].reshape(batch, nx, ny).astype(np.complex128), aclfft.BACKWARD) / (nx * ny)
    assert oracle.rel_l2(y.cpu().numpy().reshape(batch, nx, ny), ref) <= fp32_tol(nx * ny)
    p.destroy()
    v = crandn((1, nx, ny), 10).astype(np.complex64)
    vt = torch.from_numpy(v).cuda()
    q = aclfft.Plan.plan3d(1, nx, ny, aclfft.C2C)
    assert q.describe().startswith("fused2d")
    q.exec_c2c(vt, vt, aclfft.FORWARD)
    torch.cuda.synchronize()
    assert oracle.rel_l2(vt.cpu().numpy(), oracle.c2c_2d(v.astype(np.complex128))) <= fp32_tol(nx * ny)
    q.destroy()


@pytest.mark.parametrize("nx,ny,batch,flags", [(128, 128, 1024, "ntstore"), (128, 128, 6144, "ntload ntstore"),
                                               (32, 32, 1024, "")])
def test_fused_2d_streaming(cuda, monkeypatch, nx, ny, batch, flags):
    """Streaming stores for >= 16 MiB of output, streaming loads too for >= 384 MiB of input plus
    output (the plan's description names them): the first and last images against the FP64
    oracle, and the whole output bit-identical to cached accesses (ACLFFT_STREAM_STORES=0)."""
    import torch
    g = torch.Generator(device=cuda).manual_seed(nx + batch)
    x = torch.randn((batch, nx, ny), dtype=torch.complex64, device=cuda, generator=g)
    outs = []
    for rule in ("1", "0"):
        monkeypatch.setenv("ACLFFT_STREAM_STORES", rule)
        p = aclfft.Plan.plan2d(batch, nx, ny, aclfft.C2C)
        d = p.describe()
        want = flags if rule == "1" else ""
        assert d.endswith(want), d
        assert ("ntload" in d) == ("ntload" in want) and ("ntstore" in d) == ("ntstore" in want), d
        y = torch.empty_like(x)
        p.exec_c2c(x, y, aclfft.FORWARD)
        torch.cuda.synchronize()
        p.destroy()
        outs.append(y)
    assert torch.equal(outs[0], outs[1])
    for sl in (slice(0, 4), slice(batch - 4, batch)):
        ref = oracle.c2c_2d(x[sl].cpu().numpy().astype(np.complex128))
        assert oracle.rel_l2(outs[0][sl].cpu().numpy(), ref) <= fp32_tol(nx * ny)


@pytest.mark.parametrize("nx,ny,batch", [(128, 128, 2048 + 3), (64, 128, 4096 + 3), (128, 64, 4096 + 3),
                                         (64, 64, 2048 + 3), (32, 128, 2048 + 3), (32, 64, 4096 + 3)])
def test_fused_2d_pipelined(cuda, monkeypatch, nx, ny, batch):
    """The pipelined one-pass kernel (resident workgroups walking several images each, the next
    image's loads in flight through the passes; sizes with rows of 64 or more, from 2 images per
    resident workgroup: 256 of 128^2, 512 of 64 x 128, 1024 of 64^2 on this GPU) against the
    one-image-per-workgroup form (ACLFFT_2D_PIPE=0): bit-identical, with a ragged last round; a
    sample of images against the FP64 oracle."""
    import torch
    g = torch.Generator(device=cuda).manual_seed(nx * ny)
    x = torch.randn((batch, nx, ny), dtype=torch.complex64, device=cuda, generator=g)
    outs = []
    for pipe in ("1", "0"):
        monkeypatch.setenv("ACLFFT_2D_PIPE", pipe)
        p = aclfft.Plan.plan2d(batch, nx, ny, aclfft.C2C)
        y = torch.empty_like(x)
        p.exec_c2c(x, y, aclfft.BACKWARD)
        torch.cuda.synchronize()
        p.destroy()
        outs.append(y)
    assert torch.equal(outs[0], outs[1])
    for sl in (slice(0, 3), slice(1000, 1003), slice(batch - 3, batch)):
        ref = oracle.c2c_2d(x[sl].cpu().numpy().astype(np.complex128), aclfft.BACKWARD)
        assert oracle.rel_l2(outs[0][sl].cpu().numpy(), ref) <= fp32_tol(nx * ny)
