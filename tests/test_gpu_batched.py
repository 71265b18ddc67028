"""GPU: uneven-length batched plans (aclfftExtPlanManyUneven) - the persistent kernel
(ACLFFT_UNEVEN_MODE=1; the default for several lengths) and per-length tile launches
(ACLFFT_UNEVEN_MODE=2; the default for one length, or >= 2^24 points per length), each replayed
from a hipGraph."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

LENGTHS = (60, 128, 243, 1000, 1024, 2048, 4096, 6561)


def uneven_workload(count, seed, lengths=LENGTHS, max_batch=64):
    """lengths drawn log-uniformly from the set (uniform over log n, then the nearest member),
    batches uniform in [1, max_batch] - the uneven-batch benchmark configuration."""
    rng = np.random.default_rng(seed)
    logs = np.log(np.array(lengths, float))
    u = rng.uniform(logs.min(), logs.max(), count)
    ln = [int(lengths[int(np.argmin(np.abs(logs - v)))]) for v in u]
    bt = [int(b) for b in rng.integers(1, max_batch + 1, count)]
    return ln, bt


def check_groups(out, x, groups, direction):
    off = 0
    for n, b in groups:
        seg_o = out[off:off + n * b].reshape(b, n)
        seg_x = x[off:off + n * b].reshape(b, n).astype(np.complex128)
        err = oracle.rel_l2(seg_o, oracle.c2c(seg_x, direction))
        assert err <= fp32_tol(n), (n, b, err)
        off += n * b


# (ACLFFT_UNEVEN_MODE, ACLFFT_PERSISTENT_SCHEDULE, ACLFFT_UNEVEN_LANES): persistent dynamic,
# persistent static, per-length on the plan's stream, per-length over three side streams
FORMS = [(1, 0, 0), (1, 1, 0), (2, 1, 0), (2, 1, 3)]


@pytest.mark.parametrize("mode,sched,lanes", FORMS)
@pytest.mark.parametrize("graph", [True, False])
def test_uneven_plan_vs_oracle(cuda, monkeypatch, graph, mode, sched, lanes):
    import torch
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", str(mode))
    monkeypatch.setenv("ACLFFT_PERSISTENT_SCHEDULE", str(sched))
    monkeypatch.setenv("ACLFFT_UNEVEN_LANES", str(lanes))
    ln, bt = uneven_workload(40, 7)
    groups = list(zip(ln, bt))
    total = sum(n * b for n, b in groups)
    rng = np.random.default_rng(1)
    x = (rng.uniform(-1, 1, total) + 1j * rng.uniform(-1, 1, total)).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    p.set_graph_mode(graph)
    print(p.describe())
    for direction in (aclfft.FORWARD, aclfft.BACKWARD):
        for rep in range(3):  # graph replays must all be exact
            yt.zero_()
            p.exec_c2c(xt, yt, direction)
            torch.cuda.synchronize()
            out = yt.cpu().numpy()
            check_groups(out, x, groups, direction)
    p.destroy()


@pytest.mark.parametrize("stream", [0, 1])
@pytest.mark.parametrize("cset", [0, 1])
@pytest.mark.parametrize("sched", [0, 1])
def test_every_class_set(cuda, monkeypatch, cset, sched, stream):
    """Each compiled class set of the persistent kernel (ACLFFT_PERSISTENT_CLASS_SET), under both
    schedules, with cached and streaming accesses (ACLFFT_PERSISTENT_STREAM), on the benchmark's
    length mix: every group against the FP64 oracle."""
    import torch
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", "1")
    monkeypatch.setenv("ACLFFT_PERSISTENT_SCHEDULE", str(sched))
    monkeypatch.setenv("ACLFFT_PERSISTENT_CLASS_SET", str(cset))
    monkeypatch.setenv("ACLFFT_PERSISTENT_STREAM", str(stream))
    ln, bt = uneven_workload(48, 11 + cset)
    groups = list(zip(ln, bt))
    total = sum(n * b for n, b in groups)
    x = crandn(total, 5 + cset).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    d = p.describe()
    print(d)
    assert "set=%d sched=%d" % (cset, sched) in d
    assert ("stream" in d) == bool(stream)
    for direction in (aclfft.FORWARD, aclfft.BACKWARD):
        p.exec_c2c(xt, yt, direction)
        torch.cuda.synchronize()
        check_groups(yt.cpu().numpy(), x, groups, direction)
    p.destroy()


@pytest.mark.parametrize("mode,sched,lanes", FORMS)
def test_uneven_plan_mixed_with_generic_lengths(cuda, monkeypatch, mode, sched, lanes):
    """Lengths without a compiled class run as extra launches inside the same plan/graph."""
    import torch
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", str(mode))
    monkeypatch.setenv("ACLFFT_PERSISTENT_SCHEDULE", str(sched))
    monkeypatch.setenv("ACLFFT_UNEVEN_LANES", str(lanes))
    groups = [(60, 3), (420, 5), (1024, 7), (46200, 1), (7, 9), (6561, 2)]
    ln = [g[0] for g in groups]
    bt = [g[1] for g in groups]
    total = sum(n * b for n, b in groups)
    x = crandn(total, 3).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    print(p.describe())
    p.exec_c2c(xt, xt, aclfft.FORWARD)  # in place
    torch.cuda.synchronize()
    check_groups(xt.cpu().numpy(), x, groups, aclfft.FORWARD)
    p.destroy()


def _tile_variants():
    seen, out = {}, []
    for k in aclfft.kernel_registry():
        if k["kind"] != 7:
            continue
        idx = seen.get(k["n"], 0)
        seen[k["n"]] = idx + 1
        out.append((k["n"], idx))
    return out


@pytest.mark.parametrize("n,variant", _tile_variants())
def test_every_tile_kernel(cuda, monkeypatch, n, variant):
    """Each compiled tile-list kernel, forced with ACLFFT_KERNEL_VARIANT, on groups whose batches
    leave partial tiles and odd element offsets."""
    import torch
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", "2")
    monkeypatch.setenv("ACLFFT_KERNEL_VARIANT", str(variant))
    groups = [(n, 3), (63, 1), (n, 5), (n, 1)]
    ln = [g[0] for g in groups]
    bt = [g[1] for g in groups]
    total = sum(a * b for a, b in groups)
    x = crandn(total, n + variant).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    d = p.describe()
    assert "kind7 n=%d" % n in d, d
    p.exec_c2c(xt, yt, aclfft.BACKWARD)
    torch.cuda.synchronize()
    print(d)
    check_groups(yt.cpu().numpy(), x, groups, aclfft.BACKWARD)
    p.destroy()


@pytest.mark.parametrize("sched", [0, 1])
@pytest.mark.parametrize("blocks_per_cu", [1, 3, 40])
def test_persistent_grid_sizes(cuda, monkeypatch, sched, blocks_per_cu):
    """More workgroups than tiles, and far fewer: every tile is taken exactly once either way."""
    import torch
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", "1")
    monkeypatch.setenv("ACLFFT_PERSISTENT_SCHEDULE", str(sched))
    monkeypatch.setenv("ACLFFT_PERSISTENT_BLOCKS_PER_CU", str(blocks_per_cu))
    ln, bt = uneven_workload(24, 11)
    groups = list(zip(ln, bt))
    total = sum(n * b for n, b in groups)
    x = crandn(total, 5).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    p.exec_c2c(xt, yt, aclfft.FORWARD)
    torch.cuda.synchronize()
    check_groups(yt.cpu().numpy(), x, groups, aclfft.FORWARD)
    p.destroy()


def test_default_form(cuda, monkeypatch):
    """Without ACLFFT_UNEVEN_MODE: one length, or every length with >= 2^24 points, runs as
    per-length tile launches on the plan's stream; several lengths with less work share the
    persistent kernel."""
    monkeypatch.delenv("ACLFFT_UNEVEN_MODE", raising=False)
    monkeypatch.delenv("ACLFFT_UNEVEN_LANES", raising=False)
    cases = [(([4096, 4096], [3, 5]), "per-length"),
             (([1024, 4096], [1 << 14, 1 << 12]), "per-length"),
             (([1024, 4096], [1 << 14, (1 << 12) - 1]), "persistent"),
             (([60, 1000, 6561], [2, 3, 4]), "persistent")]
    for (ln, bt), want in cases:
        p = aclfft.Plan.plan_many_uneven(ln, bt)
        d = p.describe()
        p.destroy()
        if want == "persistent":
            assert d.startswith("persistent"), (ln, bt, d)
        else:
            assert "persistent" not in d and "kind7" in d and "[lane" not in d, (ln, bt, d)


@pytest.mark.parametrize("sched", [0, 1])
@pytest.mark.parametrize("prologue", [1, 3, 1000])
def test_persistent_prologue(cuda, monkeypatch, sched, prologue):
    """The small-tile prologue (ACLFFT_PERSISTENT_PROLOGUE tiles per workgroup, capped at half the
    list; 1000 exceeds every grid here) only reorders the walk: each group against the FP64
    oracle and the output bit-identical to the walk without it."""
    import torch
    monkeypatch.setenv("ACLFFT_UNEVEN_MODE", "1")
    monkeypatch.setenv("ACLFFT_PERSISTENT_SCHEDULE", str(sched))
    ln, bt = uneven_workload(64, 13)
    groups = list(zip(ln, bt))
    total = sum(n * b for n, b in groups)
    x = crandn(total, 6).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    outs = []
    for pro in (prologue, 0):
        monkeypatch.setenv("ACLFFT_PERSISTENT_PROLOGUE", str(pro))
        p = aclfft.Plan.plan_many_uneven(ln, bt)
        assert "prologue=%d" % pro in p.describe()
        yt = torch.empty_like(xt)
        p.exec_c2c(xt, yt, aclfft.FORWARD)
        torch.cuda.synchronize()
        outs.append(yt.cpu().numpy())
        p.destroy()
    check_groups(outs[0], x, groups, aclfft.FORWARD)
    assert np.array_equal(outs[0], outs[1])
