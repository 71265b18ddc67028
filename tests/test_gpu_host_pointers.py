"""GPU: host (pageable numpy) buffers on every plan form - the reference's calling convention
(its executions take host pointers, stage them through device memory and return synchronously,
lib/fft_exec_api.cpp:32-85, docs/zh/API_Reference/FFT_Common_Interface.md:5) - including the
in-place real transforms whose one buffer holds batch * 2 * (n/2 + 1) floats, and mixed host /
device pointer pairs.  Results are compared with the FP64 oracle."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,batch", [(4096, 16), (3000, 8), (1 << 16, 2)])
def test_r2c_c2r_host(cuda, n, batch):
    h = n // 2 + 1
    x = np.random.default_rng(n).standard_normal((batch, n)).astype(np.float32)
    X = np.zeros((batch, h), np.complex64)
    pr = aclfft.Plan.plan1d(n, aclfft.R2C, batch)
    pr.exec_r2c(x, X)
    ref = oracle.r2c(x.astype(np.float64))
    assert oracle.rel_l2(X, ref) <= fp32_tol(n)
    # in place on one host buffer: packed real rows in, half-spectrum rows out
    buf = np.zeros(batch * 2 * h, np.float32)
    buf[: batch * n] = x.ravel()
    pr.exec_r2c(buf, buf.view(np.complex64))
    assert np.array_equal(buf.view(np.complex64).reshape(batch, h), X)
    pr.destroy()
    pc = aclfft.Plan.plan1d(n, aclfft.C2R, batch)
    y = np.zeros((batch, n), np.float32)
    pc.exec_c2r(X, y)
    assert oracle.rel_l2(y, oracle.c2r(X.astype(np.complex128), n)) <= fp32_tol(n)
    # in place: half-spectrum rows in, packed real rows out
    buf2 = X.ravel().copy()
    pc.exec_c2r(buf2, buf2.view(np.float32))
    assert np.array_equal(buf2.view(np.float32)[: batch * n].reshape(batch, n), y)
    pc.destroy()


def test_c2c_forms_host(cuda):
    # 2D, VERTICAL, four-step and 3D plans on host buffers
    cases = [("2d", (3, 256, 384)), ("vertical", (1024, 96)), ("fourstep", (2, 1 << 18)), ("3d", (16, 32, 24))]
    for kind, shape in cases:
        x = crandn(shape, sum(shape)).astype(np.complex64)
        y = np.zeros_like(x)
        if kind == "2d":
            p = aclfft.Plan.plan2d(shape[0], shape[1], shape[2], aclfft.C2C)
            ref = oracle.c2c_2d(x.astype(np.complex128))
            m = shape[1] * shape[2]
        elif kind == "vertical":
            p = aclfft.Plan.plan1d(shape[0], aclfft.C2C, shape[1], aclfft.VERTICAL)
            ref = oracle.c2c_vertical(x.astype(np.complex128))
            m = shape[0]
        elif kind == "fourstep":
            p = aclfft.Plan.plan1d(shape[1], aclfft.C2C, shape[0])
            ref = oracle.c2c(x.astype(np.complex128))
            m = shape[1]
        else:
            p = aclfft.Plan.plan3d(*shape, aclfft.C2C)
            ref = oracle.c2c_3d(x.astype(np.complex128))
            m = shape[0] * shape[1] * shape[2]
        p.exec_c2c(x, y, aclfft.FORWARD)
        err = oracle.rel_l2(y, ref)
        print(kind, p.describe(), err)
        assert err <= fp32_tol(m), (kind, err)
        z = x.copy()
        p.exec_c2c(z, z, aclfft.FORWARD)  # in place on the host buffer
        assert np.array_equal(z, y), kind
        p.destroy()


def test_mixed_host_device(cuda):
    import torch
    n, batch = 2048, 12
    x = crandn((batch, n), 9).astype(np.complex64)
    ref = oracle.c2c(x.astype(np.complex128))
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    yd = torch.zeros((batch, n), dtype=torch.complex64, device="cuda")
    p.exec_c2c(x, yd, aclfft.FORWARD)  # host in, device out
    torch.cuda.synchronize()
    assert oracle.rel_l2(yd.cpu().numpy(), ref) <= fp32_tol(n)
    yh = np.zeros_like(x)
    p.exec_c2c(torch.from_numpy(x).cuda(), yh, aclfft.FORWARD)  # device in, host out
    assert oracle.rel_l2(yh, ref) <= fp32_tol(n)
    p.destroy()
