"""GPU: the MFMA split-precision C2C path (aclfftExtSetPrecision) against the FP64 oracle.

Bars (the precision workload, C2C n=1024 with the 60 dB case): the FP16 hi/lo and BF16 hi/mid/lo paths must be no worse than
twice the FP32 Stockham path's error against FP64 and within rel-L2 1e-6; single-BF16 is reported
(about 3 significant digits) with a loose bound.
"""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn

pytestmark = pytest.mark.gpu

N = 1024


def run(torch, x, mode, direction=aclfft.FORWARD):
    xt = torch.from_numpy(x.astype(np.complex64)).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan1d(N, aclfft.C2C, x.shape[0])
    if mode:
        p.set_precision(mode)
    p.exec_c2c(xt, yt, direction)
    torch.cuda.synchronize()
    d = p.describe()
    p.destroy()
    return yt.cpu().numpy(), d


def sixty_db(batch, seed):
    """Sum of 8 random-frequency complex tones (amplitudes 1 x4, 1e-3 x4) + 1e-6 complex noise."""
    rng = np.random.default_rng(seed)
    n = np.arange(N)
    x = np.zeros((batch, N), complex)
    for b in range(batch):
        f = rng.choice(N, 8, replace=False)
        amp = np.array([1, 1, 1, 1, 1e-3, 1e-3, 1e-3, 1e-3])
        ph = rng.uniform(0, 2 * np.pi, 8)
        x[b] = (amp * np.exp(1j * (2 * np.pi * np.outer(n, f) / N + ph))).sum(1)
    return x + 1e-6 * crandn((batch, N), seed + 1)


@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_mfma_modes_vs_fp64(cuda, direction):
    import torch
    x = crandn((300, N), 11).astype(np.complex64)
    ref = oracle.c2c(x.astype(np.complex128), direction)
    y32, d32 = run(torch, x, 0, direction)
    e32 = oracle.rel_l2(y32, ref)
    errs = {}
    for mode in (aclfft.PRECISION_BF16, aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
        y, d = run(torch, x, mode, direction)
        assert d.startswith("mfma"), d
        errs[mode] = oracle.rel_l2(y, ref)
    print("rel_l2 vs FP64: fp32 %.3e  bf16 %.3e  fp16x2 %.3e  bf16x3 %.3e" % (
        e32, errs[1], errs[2], errs[3]))
    assert errs[aclfft.PRECISION_BF16] < 2e-2
    for mode in (aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
        assert errs[mode] <= max(2 * e32, 1e-7), (mode, errs[mode], e32)
        assert errs[mode] <= 1e-6


def test_mfma_sixty_db(cuda):
    import torch
    x = sixty_db(64, 3).astype(np.complex64)
    ref = oracle.c2c(x.astype(np.complex128))
    e32 = oracle.rel_l2(run(torch, x, 0)[0], ref)
    for mode in (aclfft.PRECISION_FP16X2, aclfft.PRECISION_BF16X3):
        y, _ = run(torch, x, mode)
        e = oracle.rel_l2(y, ref)
        print("60 dB: mode %d rel_l2 %.3e (fp32 %.3e)" % (mode, e, e32))
        assert e <= max(2 * e32, 1e-7)


def test_mfma_ragged_batch_and_in_place(cuda):
    import torch
    for batch in (1, 3, 5, 4099):
        x = crandn((batch, N), batch).astype(np.complex64)
        xt = torch.from_numpy(x).cuda()
        p = aclfft.Plan.plan1d(N, aclfft.C2C, batch)
        p.set_precision(aclfft.PRECISION_FP16X2)
        p.exec_c2c(xt, xt, aclfft.FORWARD)
        torch.cuda.synchronize()
        assert oracle.rel_l2(xt.cpu().numpy(), oracle.c2c(x.astype(np.complex128))) < 1e-6
        p.destroy()


def test_mfma_precision_gate(cuda):
    p = aclfft.Plan.plan1d(512, aclfft.C2C, 4)
    with pytest.raises(aclfft.AclfftError) as e:
        p.set_precision(aclfft.PRECISION_FP16X2)
    assert e.value.code == aclfft.NOT_SUPPORTED
    assert p.describe().startswith("kind0")
