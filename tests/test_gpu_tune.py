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
    hashes = {tok for line in text.splitlines() for tok in line.split() if tok.startswith("u")}
    assert len(hashes) == 2, text  # two group hashes
    assert all(" plan p0 2:" in line for line in text.splitlines() if " plan " in line), text  # the forced form


def uneven_groups(count, seed, lengths=(60, 128, 243, 1000, 1024, 2048, 4096, 6561)):
    rng = np.random.default_rng(seed)
    logs = np.log(np.array(lengths, float))
    u = rng.uniform(logs.min(), logs.max(), count)
    ln = [int(lengths[int(np.argmin(np.abs(logs - v)))]) for v in u]
    bt = [int(b) for b in rng.integers(1, 65, count)]
    return ln, bt


def check_uneven(p, ln, bt, seed):
    import torch
    groups = list(zip(ln, bt))
    total = sum(n * b for n, b in groups)
    x = crandn(total, seed).astype(np.complex64)
    y = torch.empty(total, dtype=torch.complex64, device="cuda")
    p.exec_c2c(torch.from_numpy(x).cuda(), y, aclfft.FORWARD)
    torch.cuda.synchronize()
    out = y.cpu().numpy()
    off = 0
    for n, b in groups:
        ref = oracle.c2c(x[off:off + n * b].reshape(b, n).astype(np.complex128))
        assert oracle.rel_l2(out[off:off + n * b].reshape(b, n), ref) <= fp32_tol(n), (n, b)
        off += n * b


@pytest.mark.parametrize("auto", [False, True])
def test_tuned_uneven_plan(cuda, tmp_path, monkeypatch, auto):
    """aclfftExtTune on an uneven plan: the plan-level choice (form, schedule, workgroups per CU,
    class set; with automatic precision also FP32 or an MFMA split) and the step kernels; the
    tuned plan matches the oracle, and a new plan of the same groups takes the cached choice."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("ACLFFT_TUNE_CACHE", str(cache))
    monkeypatch.delenv("ACLFFT_TUNE", raising=False)
    ln, bt = uneven_groups(60, 3)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    if auto:
        p.set_precision(aclfft.PRECISION_AUTO)
    before = p.describe()
    p.tune()
    after = p.describe()
    print("default:", before)
    print("tuned:  ", after)
    check_uneven(p, ln, bt, 5)
    text = cache.read_text()
    assert (" plan auto " if auto else " plan p0 ") in text
    q = aclfft.Plan.plan_many_uneven(ln, bt)
    if auto:
        q.set_precision(aclfft.PRECISION_AUTO)
    check_uneven(q, ln, bt, 6)  # first execution applies the cached plan choice
    assert q.describe() == after
    p.destroy()
    q.destroy()


@pytest.mark.parametrize("kind,n", [("c2c", 512), ("c2c", 1024), ("c2c", 4096), ("r2c", 8192)])
def test_automatic_precision(cuda, tmp_path, monkeypatch, kind, n):
    """ACLFFT_EXT_PRECISION_AUTO: FP32 until tuned; the tuner keeps FP32 or an MFMA split within
    1e-6 of the FP32 result (single BF16 never qualifies); the result meets the FP32 bar."""
    import torch
    monkeypatch.setenv("ACLFFT_TUNE_CACHE", str(tmp_path / "tune.txt"))
    b = 512
    t = aclfft.C2C if kind == "c2c" else aclfft.R2C
    p = aclfft.Plan.plan1d(n, t, b)
    p.set_precision(aclfft.PRECISION_AUTO)
    assert "mfma" not in p.describe()
    p.tune()
    d = p.describe()
    print(kind, n, "->", d)
    assert "mode=bf16 " not in d + " "
    if kind == "c2c":
        x = crandn((b, n), 9).astype(np.complex64)
        y = torch.empty((b, n), dtype=torch.complex64, device="cuda")
        p.exec_c2c(torch.from_numpy(x).cuda(), y, aclfft.FORWARD)
        ref = oracle.c2c(x.astype(np.complex128))
    else:
        x = np.random.default_rng(9).standard_normal((b, n)).astype(np.float32)
        y = torch.empty((b, n // 2 + 1), dtype=torch.complex64, device="cuda")
        p.exec_r2c(torch.from_numpy(x).cuda(), y)
        ref = oracle.r2c(x.astype(np.float64))
    torch.cuda.synchronize()
    assert oracle.rel_l2(y.cpu().numpy(), ref) <= fp32_tol(n)
    p.destroy()


def test_automatic_precision_gate(cuda):
    p = aclfft.Plan.plan1d(1000, aclfft.C2C, 4)
    with pytest.raises(aclfft.AclfftError) as e:
        p.set_precision(aclfft.PRECISION_AUTO)
    assert e.value.code == aclfft.NOT_SUPPORTED
    p.destroy()
