"""The autotuner (ACLFFT_TUNE=1): on a plan's first execution every step with several compiled
candidates times each of them on scratch buffers sized for the largest buffer the plan uses,
keeps the fastest, and records the row- and column-kernel choices in the tuning cache (ACLFFT_TUNE_CACHE),
which later plans of the same description read without timing.  Results must match the FP64
oracle whatever the choice."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n,batch", [("c2c", 1000, 512), ("c2c", 1 << 20, 2), ("r2c", 3000, 64),
                                          ("c2r", 4096, 32), ("2d", 2048, 64)])
def test_autotune_then_cache(cuda, tmp_path, monkeypatch, kind, n, batch):
    import torch
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("ACLFFT_TUNE_CACHE", str(cache))
    described = []
    for tune in ("1", "0"):
        monkeypatch.setenv("ACLFFT_TUNE", tune)
        if kind == "c2c":
            x = crandn((batch, n), n).astype(np.complex64)
            p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
            y = torch.empty((batch, n), dtype=torch.complex64, device="cuda")
            p.exec_c2c(torch.from_numpy(x).cuda(), y, aclfft.FORWARD)
            ref, m = oracle.c2c(x.astype(np.complex128)), n
        elif kind == "2d":
            x = crandn((2, n, batch), n).astype(np.complex64)
            p = aclfft.Plan.plan2d(2, n, batch, aclfft.C2C)
            y = torch.empty_like(torch.from_numpy(x)).cuda()
            p.exec_c2c(torch.from_numpy(x).cuda(), y, aclfft.FORWARD)
            ref, m = oracle.c2c_2d(x.astype(np.complex128)), n * batch
        elif kind == "r2c":
            x = np.random.default_rng(n).standard_normal((batch, n)).astype(np.float32)
            p = aclfft.Plan.plan1d(n, aclfft.R2C, batch)
            y = torch.empty((batch, n // 2 + 1), dtype=torch.complex64, device="cuda")
            p.exec_r2c(torch.from_numpy(x).cuda(), y)
            ref, m = oracle.r2c(x.astype(np.float64)), n
        else:
            X = crandn((batch, n // 2 + 1), n)
            X[:, 0] = X[:, 0].real
            X[:, -1] = X[:, -1].real
            p = aclfft.Plan.plan1d(n, aclfft.C2R, batch)
            y = torch.empty((batch, n), dtype=torch.float32, device="cuda")
            p.exec_c2r(torch.from_numpy(X.astype(np.complex64)).cuda(), y)
            ref, m = oracle.c2r(X, n), n
        torch.cuda.synchronize()
        err = oracle.rel_l2(y.cpu().numpy(), ref)
        described.append(p.describe())
        print("tune=%s %s rel_l2 %.3e" % (tune, p.describe(), err))
        p.destroy()
        assert err <= fp32_tol(m), err
    # every tuned step (row and column kernels) recorded, then read back by the untimed plan
    assert cache.exists() and cache.read_text().strip()
    assert described[0] == described[1]


def test_autotune_streaming_variant_recorded(cuda, tmp_path, monkeypatch):
    """A working set beyond the last-level cache has streaming and ordinary variants of the same
    configuration among its candidates; the cache tells them apart, so the untimed plan runs the
    variant the tuner chose.  Rows spread over the batch are checked against the oracle."""
    import torch
    n, batch = 1024, 32768
    monkeypatch.setenv("ACLFFT_TUNE_CACHE", str(tmp_path / "tune.txt"))
    x = torch.randn(batch, n, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    rows = [0, 1, 4097, batch - 1]
    xr = x[rows].cpu().numpy().astype(np.complex128)
    described = []
    for tune in ("1", "0"):
        monkeypatch.setenv("ACLFFT_TUNE", tune)
        p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
        y.zero_()
        p.exec_c2c(x, y, aclfft.FORWARD)
        torch.cuda.synchronize()
        described.append(p.describe())
        p.destroy()
        err = oracle.rel_l2(y[rows].cpu().numpy(), oracle.c2c(xr))
        print("tune=%s %s rel_l2 %.3e" % (tune, described[-1], err))
        assert err <= fp32_tol(n), err
    assert described[0] == described[1]


def test_autotune_uneven_per_length(cuda, tmp_path, monkeypatch):
    """Per-length tile launches of an uneven plan are tuned and recorded under a key that includes
    the groups: a plan of other groups with the same lengths does not take them over blindly, and
    both still match the oracle."""
    import torch
    monkeypatch.setenv("ACLFFT_TUNE_CACHE", str(tmp_path / "tune.txt"))
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", "2")
    for groups in ([(4096, 7), (1024, 9), (4096, 2)], [(4096, 3), (1024, 1)]):
        ln = [g[0] for g in groups]
        bt = [g[1] for g in groups]
        total = sum(a * b for a, b in groups)
        x = crandn(total, total).astype(np.complex64)
        described = []
        for tune in ("1", "0"):
            monkeypatch.setenv("ACLFFT_TUNE", tune)
            p = aclfft.Plan.plan_many_uneven(ln, bt)
            yt = torch.empty(total, dtype=torch.complex64, device="cuda")
            p.exec_c2c(torch.from_numpy(x).cuda(), yt, aclfft.FORWARD)
            torch.cuda.synchronize()
            described.append(p.describe())
            p.destroy()
            out = yt.cpu().numpy()
            off = 0
            for n, b in groups:
                err = oracle.rel_l2(out[off:off + n * b].reshape(b, n),
                                    oracle.c2c(x[off:off + n * b].reshape(b, n).astype(np.complex128)))
                assert err <= fp32_tol(n), (groups, n, err)
                off += n * b
        assert described[0] == described[1], described
    text = (tmp_path / "tune.txt").read_text()
    assert len({line.split()[-2] for line in text.splitlines()}) == 2, text  # two group hashes
