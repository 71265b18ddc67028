"""GPU: uneven-length batched plans (aclfftExtPlanManyUneven) - persistent kernel + hipGraph."""
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


@pytest.mark.parametrize("graph", [True, False])
def test_uneven_plan_vs_oracle(cuda, graph):
    import torch
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


def test_uneven_plan_mixed_with_generic_lengths(cuda):
    """Lengths without a compiled class run as extra launches inside the same plan/graph."""
    import torch
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
