"""GPU: aclfftExtExecC2CNoAlloc - the C2C execution that allocates nothing (device tables and
workspace in the library's device scratch area) - against aclfftExecC2C on the same plans, called
through the C-ABI: bit-identical results (same kernels, same tables) and the FP32 bar against the FP64
oracle, on two calls in a row with the second input smaller, a four-step length whose workspace
lives in the scratch area, both directions and in place; the size limit and the pointer rule."""
import ctypes

import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

lib = aclfft.lib
lib.aclfftExtExecC2CNoAlloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
lib.aclfftExtExecC2CNoAlloc.restype = ctypes.c_int


def no_alloc(p, xin, yout, direction):
    import torch
    rc = lib.aclfftExtExecC2CNoAlloc(p.handle, xin.data_ptr(), yout.data_ptr(), direction)
    torch.cuda.synchronize()
    return rc


@pytest.fixture
def private_cache(tmp_path, monkeypatch):
    # both entry points then run the plan's default kernels (no tuned choice from a cache)
    monkeypatch.setenv("ACLFFT_TUNE_CACHE", str(tmp_path / "tune.txt"))
    monkeypatch.delenv("ACLFFT_TUNE", raising=False)


def check_case(torch, n, batch, direction, seed, in_place=False):
    x = crandn((batch, n), seed).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    p.set_stream(torch.cuda.current_stream())
    ref_dev = torch.empty_like(xt)
    p.exec_c2c(xt, ref_dev, direction)  # the existing entry point, plan-owned tables / workspace
    torch.cuda.synchronize()
    if in_place:
        y = xt.clone()
        assert no_alloc(p, y, y, direction) == aclfft.SUCCESS
    else:
        y = torch.full_like(xt, float("nan"))
        assert no_alloc(p, xt, y, direction) == aclfft.SUCCESS
    assert torch.equal(y, ref_dev), (n, batch, direction)
    err = oracle.rel_l2(y.cpu().numpy(), oracle.c2c(x.astype(np.complex128), direction))
    assert err <= fp32_tol(n), (n, batch, err)
    p.destroy()
    print("n=%d batch=%d dir=%d in_place=%s: identical to aclfftExecC2C, rel-L2 %.3e" % (
        n, batch, direction, in_place, err), flush=True)


def test_two_calls_second_smaller(cuda, private_cache):
    import torch
    # 2^20 x 4: four-step (32 MiB of workspace plus the FP64 tables in the scratch area), then a
    # single-kernel length with a smaller input: the second call finds the first call's contents
    check_case(torch, 1 << 20, 4, aclfft.FORWARD, 1)
    check_case(torch, 1024, 16, aclfft.FORWARD, 2)


@pytest.mark.parametrize("n,batch", [(1000, 7), (6561, 3), (46200, 2), (1 << 18, 3)])
def test_lengths_and_directions(cuda, private_cache, n, batch):
    import torch
    check_case(torch, n, batch, aclfft.BACKWARD, n)
    check_case(torch, n, batch, aclfft.FORWARD, n + 1, in_place=True)


def test_limits(cuda, private_cache):
    import torch
    # 2^22 x 4 needs 128 MiB of workspace: more than the scratch area
    p = aclfft.Plan.plan1d(1 << 22, aclfft.C2C, 4)
    x = torch.zeros(4 << 22, dtype=torch.complex64, device="cuda")
    assert no_alloc(p, x, x, aclfft.FORWARD) == aclfft.INVALID_SIZE
    p.destroy()
    # host pointers are refused
    q = aclfft.Plan.plan1d(64, aclfft.C2C, 2)
    h = np.zeros((2, 64), np.complex64)
    rc = lib.aclfftExtExecC2CNoAlloc(q.handle, h.ctypes.data, h.ctypes.data, aclfft.FORWARD)
    assert rc == aclfft.INVALID_VALUE
    # a real plan is the wrong type
    r = aclfft.Plan.plan1d(64, aclfft.R2C, 2)
    d = torch.zeros(2 * 64, device="cuda")
    assert lib.aclfftExtExecC2CNoAlloc(r.handle, d.data_ptr(), d.data_ptr(), aclfft.FORWARD) == aclfft.INVALID_TYPE
    q.destroy()
    r.destroy()
