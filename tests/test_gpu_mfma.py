"""GPU: the MFMA split-precision C2C path (aclfftExtSetPrecision) against the FP64 oracle.

Every MFMA kernel of the registry (n = r1 * r2 with 16-, 32- and 64-point dense DFT stages:
256 = 16x16, 512 = 16x32 / 32x16, 1024 = 32x32 on both MFMA tiles / 16x64 / 64x16, 2048 = 32x64 /
64x32, 4096 = 64x64; and n = 16, 32, 64 as one dense stage), forced one by one with
ACLFFT_MFMA_VARIANT, in every operand precision; 2D images as F X F.

Bars (the precision workload): the FP16 hi/lo and BF16 hi/mid/lo paths must be no worse than twice
the FP32 Stockham path's error against FP64 and within rel-L2 1e-6 * log2(n); single-BF16 is
reported (about 3 significant digits) with a loose bound.  The 60 dB dynamic-range case (8 tones
at amplitudes 1 and 1e-3 plus 1e-6 noise) is held to the same bars.
"""
import ctypes

import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

MODES = (aclfft.PRECISION_BF16, aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3)


class MfmaInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("r1", ctypes.c_int), ("r2", ctypes.c_int), ("shape", ctypes.c_int),
                ("tiled", ctypes.c_int)]


def mfma_variants():
    """(n, variant index among the entries of length n, r1, r2, shape) of every registry entry."""
    lib = aclfft.lib
    out, seen = [], {}
    for i in range(lib.ofx_mfma_kernel_count()):
        mi = MfmaInfo()
        assert lib.ofx_mfma_kernel_info_get(i, ctypes.byref(mi)) == 0
        k = seen.get(mi.n, 0)
        seen[mi.n] = k + 1
        out.append((mi.n, k, mi.r1, mi.r2, mi.shape))
    return out


VARIANTS = mfma_variants()


def run(torch, x, mode, direction=aclfft.FORWARD):
    xt = torch.from_numpy(x.astype(np.complex64)).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan1d(x.shape[1], aclfft.C2C, x.shape[0])
    if mode:
        p.set_precision(mode)
    p.exec_c2c(xt, yt, direction)
    torch.cuda.synchronize()
    d = p.describe()
    p.destroy()
    return yt.cpu().numpy(), d


def sixty_db(batch, n, seed):
    """Sum of 8 random-frequency complex tones (amplitudes 1 x4, 1e-3 x4) + 1e-6 complex noise."""
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    x = np.zeros((batch, n), complex)
    for b in range(batch):
        f = rng.choice(n, 8, replace=False)
        amp = np.array([1, 1, 1, 1, 1e-3, 1e-3, 1e-3, 1e-3])
        ph = rng.uniform(0, 2 * np.pi, 8)
        x[b] = (amp * np.exp(1j * (2 * np.pi * np.outer(t, f) / n + ph))).sum(1)
    return x + 1e-6 * crandn((batch, n), seed + 1)


def test_registry_covers_the_reference_stage_shapes():
    shapes = {(n, r1, r2) for n, _, r1, r2, _ in VARIANTS}
    # the radix pairs of the reference's matrix-DFT kernel (fft_b/fft1_d_b.cpp:27-48), in both orders
    for n, a, b in ((256, 16, 16), (512, 16, 32), (1024, 16, 64), (2048, 32, 64), (4096, 64, 64)):
        assert (n, a, b) in shapes or (n, b, a) in shapes, (n, a, b)


@pytest.mark.parametrize("n,variant,r1,r2,shape", VARIANTS)
@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_mfma_kernels_vs_fp64(cuda, monkeypatch, n, variant, r1, r2, shape, direction):
    import torch
    monkeypatch.setenv("ACLFFT_MFMA_VARIANT", str(variant))
    x = crandn((37, n), 11 + n).astype(np.complex64)
    ref = oracle.c2c(x.astype(np.complex128), direction)
    y32, _ = run(torch, x, 0, direction)
    e32 = oracle.rel_l2(y32, ref)
    errs = {}
    for mode in MODES:
        y, d = run(torch, x, mode, direction)
        assert d.startswith("mfma n=%d %dx%d shape=%d" % (n, r1, r2, shape)), d
        errs[mode] = oracle.rel_l2(y, ref)
    print("n=%d %dx%d/%d rel_l2 vs FP64: fp32 %.3e  bf16 %.3e  fp16x2 %.3e  bf16x3 %.3e" % (
        n, r1, r2, shape, e32, errs[1], errs[2], errs[3]))
    assert errs[aclfft.PRECISION_BF16] < 2e-2
    for mode in (aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
        assert errs[mode] <= max(2 * e32, 1e-7), (mode, errs[mode], e32)
        assert errs[mode] <= fp32_tol(n)


@pytest.mark.parametrize("n", sorted({v[0] for v in VARIANTS}))
def test_mfma_sixty_db(cuda, n):
    import torch
    x = sixty_db(16, n, 3).astype(np.complex64)
    ref = oracle.c2c(x.astype(np.complex128))
    e32 = oracle.rel_l2(run(torch, x, 0)[0], ref)
    for mode in (aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
        y, _ = run(torch, x, mode)
        e = oracle.rel_l2(y, ref)
        print("60 dB n=%d: mode %d rel_l2 %.3e (fp32 %.3e)" % (n, mode, e, e32))
        assert e <= max(2 * e32, 1e-7)


@pytest.mark.parametrize("n", [32, 256, 1024, 4096])
def test_mfma_ragged_batch_and_in_place(cuda, n):
    import torch
    for batch in (1, 3, 5, 4099):
        x = crandn((batch, n), batch).astype(np.complex64)
        xt = torch.from_numpy(x).cuda()
        p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
        p.set_precision(aclfft.PRECISION_FP16X2)
        p.exec_c2c(xt, xt, aclfft.FORWARD)
        torch.cuda.synchronize()
        assert oracle.rel_l2(xt.cpu().numpy(), oracle.c2c(x.astype(np.complex128))) < fp32_tol(n)
        p.destroy()


@pytest.mark.parametrize("n", [2048, 64])
def test_mfma_extreme_magnitudes(cuda, n):
    """FP16 operands are formed from data scaled per transform by a power of two: transforms of
    magnitude 1e-30 and 1e30 next to ordinary ones keep their accuracy; zero transforms stay zero
    (n = 64: the dense kernel, whose 16 transforms per wave each take their own scale)."""
    import torch
    x = crandn((6, n), 41)
    x[1] *= 1e-30
    x[2] *= 1e30
    x[3] = 0
    x[4] *= 1e-20
    x = x.astype(np.complex64)
    ref = oracle.c2c(x.astype(np.complex128))
    for mode in (aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
        y, _ = run(torch, x, mode)
        assert np.all(y[3] == 0)
        for b in (0, 1, 2, 4, 5):
            assert oracle.rel_l2(y[b], ref[b]) <= fp32_tol(n), (mode, b)


@pytest.mark.parametrize("n", [32, 128, 512, 2048, 4096, 8192])
def test_mfma_real_transforms(cuda, n):
    """R2C / C2R in an MFMA precision: the n/2-point MFMA transform with the Hermitian unpack /
    pre-pack kernels."""
    import torch
    b = 9
    rng = np.random.default_rng(n)
    x = rng.standard_normal((b, n)).astype(np.float32)
    xt = torch.from_numpy(x).cuda()
    X = torch.empty(b, n // 2 + 1, dtype=torch.complex64, device="cuda")
    back = torch.empty_like(xt)
    ref = oracle.r2c(x.astype(np.float64))
    for mode in (aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
        pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
        pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
        pr.set_precision(mode)
        pc.set_precision(mode)
        assert "mfma" in pr.describe() and "mfma" in pc.describe()
        pr.exec_r2c(xt, X)
        pc.exec_c2r(X, back)
        torch.cuda.synchronize()
        Xh = X.cpu().numpy()
        assert oracle.rel_l2(Xh, ref) <= fp32_tol(n)
        assert oracle.rel_l2(back.cpu().numpy(), oracle.c2r(Xh.astype(np.complex128), n)) <= fp32_tol(n)
        pr.destroy()
        pc.destroy()


def mfma_2d_count(nx, ny):
    """Number of registry kernels (entries >= 1) that run nx x ny images: r2 = nx, r1 = ny."""
    lib = aclfft.lib
    c = 0
    for i in range(1, lib.ofx_mfma_kernel_count()):
        mi = MfmaInfo()
        assert lib.ofx_mfma_kernel_info_get(i, ctypes.byref(mi)) == 0
        c += mi.r1 == ny and mi.r2 == nx
    return c


def run_2d(torch, x, mode, direction=aclfft.FORWARD, in_place=False):
    xt = torch.from_numpy(x.astype(np.complex64)).cuda()
    yt = xt if in_place else torch.empty_like(xt)
    p = aclfft.Plan.plan2d(x.shape[0], x.shape[1], x.shape[2], aclfft.C2C)
    if mode:
        p.set_precision(mode)
    p.exec_c2c(xt, yt, direction)
    torch.cuda.synchronize()
    d = p.describe()
    p.destroy()
    return yt.cpu().numpy(), d


@pytest.mark.parametrize("nx", [16, 32, 64])
@pytest.mark.parametrize("ny", [16, 32, 64])
@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_mfma_2d_vs_fp64(cuda, monkeypatch, nx, ny, direction):
    """2D C2C in an MFMA precision: each nx x ny image as F_nx X F_ny, the column and row DFTs as
    the two dense stages of an MFMA kernel (the reference's P X Q form of its 2D kernel,
    src/fft2_d/arch32/dd/fft2_d_dd.cpp:31-66), every kernel that runs the shape, every operand
    precision, against the FP64 oracle with the 1D bars at n = nx * ny."""
    import torch
    x = crandn((7, nx, ny), 5 * nx + ny + direction).astype(np.complex64)
    ref = oracle.c2c_2d(x.astype(np.complex128), direction)
    e32 = oracle.rel_l2(run_2d(torch, x, 0, direction)[0], ref)
    count = mfma_2d_count(nx, ny)
    assert count >= 1
    for variant in range(count):
        monkeypatch.setenv("ACLFFT_MFMA_VARIANT", str(variant))
        errs = {}
        for mode in MODES:
            y, d = run_2d(torch, x, mode, direction)
            assert d.startswith("mfma n=%d %dx%d" % (nx * ny, ny, nx)) and "2d=%dx%d" % (nx, ny) in d, d
            errs[mode] = oracle.rel_l2(y, ref)
        print("2D %dx%d variant %d rel_l2 vs FP64: fp32 %.3e  bf16 %.3e  fp16x2 %.3e  bf16x3 %.3e" % (
            nx, ny, variant, e32, errs[1], errs[2], errs[3]))
        assert errs[aclfft.PRECISION_BF16] < 2e-2
        for mode in (aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
            assert errs[mode] <= max(2 * e32, 1e-7), (mode, errs[mode], e32)
            assert errs[mode] <= fp32_tol(nx * ny)


def test_mfma_2d_batches_in_place_and_sixty_db(cuda):
    """Ragged image counts, in-place execution, an output that is not 16-byte aligned, and the
    60 dB tone case laid out as 32 x 32 images (a 1024-point tone signal per image), FP16 hi/lo
    and BF16 hi/mid/lo."""
    import torch
    for batch in (1, 3, 4099):
        x = crandn((batch, 32, 64), batch).astype(np.complex64)
        y, _ = run_2d(torch, x, aclfft.PRECISION_FP16X2, in_place=True)
        assert oracle.rel_l2(y, oracle.c2c_2d(x.astype(np.complex128))) <= fp32_tol(2048)
    # an output 8 bytes past a 16-byte boundary (the 2D stores fall back to 8-byte accesses)
    x = crandn((5, 64, 32), 9).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    buf = torch.zeros(x.size + 2, dtype=torch.complex64, device="cuda")
    yt = buf[1:1 + x.size].view(5, 64, 32)
    p = aclfft.Plan.plan2d(5, 64, 32, aclfft.C2C)
    p.set_precision(aclfft.PRECISION_BF16X3)
    p.exec_c2c(xt, yt, aclfft.FORWARD)
    torch.cuda.synchronize()
    p.destroy()
    assert oracle.rel_l2(yt.cpu().numpy(), oracle.c2c_2d(x.astype(np.complex128))) <= fp32_tol(2048)
    assert buf[0].item() == 0 and buf[-1].item() == 0
    x = sixty_db(16, 1024, 7).reshape(16, 32, 32).astype(np.complex64)
    ref = oracle.c2c_2d(x.astype(np.complex128))
    e32 = oracle.rel_l2(run_2d(torch, x, 0)[0], ref)
    for mode in (aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
        e = oracle.rel_l2(run_2d(torch, x, mode)[0], ref)
        print("2D 60 dB 32x32: mode %d rel_l2 %.3e (fp32 %.3e)" % (mode, e, e32))
        assert e <= max(2 * e32, 1e-7)


@pytest.mark.parametrize("rank,shape", [(2, (2, 128, 128)), (2, (2, 128, 32)), (2, (2, 48, 32)), (3, (32, 32, 32))])
def test_mfma_2d_gate(cuda, rank, shape):
    """Shapes without an MFMA form (an axis of 128 or of a length other than 16, 32, 64; 3D)
    refuse an MFMA precision and keep their FP32 launch list."""
    if rank == 2:
        p = aclfft.Plan.plan2d(*shape, aclfft.C2C)
    else:
        p = aclfft.Plan.plan3d(*shape, aclfft.C2C)
    before = p.describe()
    with pytest.raises(aclfft.AclfftError) as e:
        p.set_precision(aclfft.PRECISION_FP16X2)
    assert e.value.code == aclfft.NOT_SUPPORTED
    assert p.describe() == before
    p.destroy()


def test_mfma_dense_lengths_in_uneven_plan(cuda):
    """Uneven plans in an MFMA precision keep the dense small-n lengths on the FP32 kernels (the
    dense kernels take no tile lists) and run the rest on MFMA; every group against the oracle."""
    import torch
    ln, bt = [16, 1024, 64, 32, 1024, 16], [5, 3, 40, 17, 1, 2]
    x = crandn((sum(a * b for a, b in zip(ln, bt)),), 77).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    p.set_precision(aclfft.PRECISION_FP16X2)
    assert "mfma n=1024" in p.describe()
    p.exec_c2c(xt, yt, aclfft.FORWARD)
    torch.cuda.synchronize()
    y = yt.cpu().numpy()
    off = 0
    for a, b in zip(ln, bt):
        seg = x[off:off + a * b].reshape(b, a).astype(np.complex128)
        assert oracle.rel_l2(y[off:off + a * b].reshape(b, a), oracle.c2c(seg)) <= fp32_tol(a), (a, b)
        off += a * b
    p.destroy()


def test_mfma_precision_gate(cuda):
    p = aclfft.Plan.plan1d(1000, aclfft.C2C, 4)
    with pytest.raises(aclfft.AclfftError) as e:
        p.set_precision(aclfft.PRECISION_FP16X2)
    assert e.value.code == aclfft.NOT_SUPPORTED
    assert p.describe().startswith("kind0")
    p.destroy()


def uneven_workload(count, seed, lengths=(60, 128, 243, 256, 512, 1000, 1024, 2048, 4096, 6561)):
    rng = np.random.default_rng(seed)
    logs = np.log(np.array(lengths, float))
    u = rng.uniform(logs.min(), logs.max(), count)
    ln = [int(lengths[int(np.argmin(np.abs(logs - v)))]) for v in u]
    bt = [int(b) for b in rng.integers(1, 40, count)]
    return ln, bt


@pytest.mark.parametrize("mode", [aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3])
@pytest.mark.parametrize("graph", [False, True])
def test_mfma_uneven_plan(cuda, mode, graph):
    """Uneven plans in an MFMA precision: lengths with MFMA kernels run as MFMA tile launches,
    the others on the FP32 kernels; every group against the FP64 oracle, both directions, with an
    output scale, odd group offsets included ((243, b) groups shift the following ones by 8 B)."""
    import torch
    ln, bt = uneven_workload(30, 7)
    ln += [243, 1024, 7, 4096]
    bt += [1, 3, 2, 5]
    groups = list(zip(ln, bt))
    total = sum(n * b for n, b in groups)
    x = crandn(total, 8).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    p.set_precision(mode)
    p.set_graph_mode(graph)
    p.set_scale(0.5)
    d = p.describe()
    print(d)
    assert "mfma" in d
    for direction in (aclfft.FORWARD, aclfft.BACKWARD):
        for rep in range(2):
            yt.zero_()
            p.exec_c2c(xt, yt, direction)
            torch.cuda.synchronize()
            out = yt.cpu().numpy()
            off = 0
            for n, b in groups:
                ref = 0.5 * oracle.c2c(x[off:off + n * b].reshape(b, n).astype(np.complex128), direction)
                err = oracle.rel_l2(out[off:off + n * b].reshape(b, n), ref)
                assert err <= fp32_tol(n), (n, b, err)
                off += n * b
    p.destroy()


def test_mfma_uneven_gate(cuda):
    p = aclfft.Plan.plan_many_uneven([60, 243, 1000], [2, 3, 4])
    with pytest.raises(aclfft.AclfftError) as e:
        p.set_precision(aclfft.PRECISION_FP16X2)
    assert e.value.code == aclfft.NOT_SUPPORTED
    assert p.describe().startswith("persistent")
    p.destroy()
