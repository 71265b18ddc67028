"""GPU parity of the double-precision transforms (aclfftExecZ2Z / D2Z / Z2D) against the FP64
oracle.  Tolerance: rel-L2 <= 1e-14 * log2(n) (FP32 arithmetic would give ~1e-7), both
directions, batches that leave partial workgroups, host and device pointers, in place, adjoints."""
import math

import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn

pytestmark = pytest.mark.gpu


def fp64_tol(n):
    return 1e-14 * max(1.0, math.log2(max(n, 2)))


def _lengths(kind):
    return sorted({k["n"] for k in aclfft.kernel_registry() if k["kind"] == kind})


Z2Z_N = _lengths(8)
REAL_EVEN_N = [2 * m for m in _lengths(9)]
REAL_ODD_N = _lengths(11)


@pytest.mark.parametrize("n", Z2Z_N)
@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_z2z_vs_oracle(cuda, n, direction):
    import torch
    batch = 37
    x = crandn((batch, n), seed=n + 7)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p = aclfft.Plan.plan1d(n, aclfft.Z2Z, batch)
    p.exec_z2z(xt, yt, direction)
    torch.cuda.synchronize()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c(x, direction))
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    assert err <= fp64_tol(n), err


@pytest.mark.parametrize("n", REAL_EVEN_N + REAL_ODD_N)
def test_d2z_z2d_vs_oracle(cuda, n):
    import torch
    batch = 19
    h = n // 2 + 1
    x = np.random.default_rng(n).standard_normal((batch, n))
    xt = torch.from_numpy(x).cuda()
    Xt = torch.empty((batch, h), dtype=torch.complex128, device="cuda")
    p = aclfft.Plan.plan1d(n, aclfft.D2Z, batch)
    p.exec_d2z(xt, Xt)
    torch.cuda.synchronize()
    err = oracle.rel_l2(Xt.cpu().numpy(), oracle.r2c(x))
    assert err <= fp64_tol(n), ("d2z", err)
    X = crandn((batch, h), n + 1)
    X[:, 0] = X[:, 0].real
    if n % 2 == 0:
        X[:, -1] = X[:, -1].real
    q = aclfft.Plan.plan1d(n, aclfft.Z2D, batch)
    yt = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    q.exec_z2d(torch.from_numpy(X).cuda(), yt)
    torch.cuda.synchronize()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2r(X, n))
    print(p.describe(), "|", q.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    q.destroy()
    assert err <= fp64_tol(n), ("z2d", err)


def test_z2z_host_pointers_and_in_place(cuda):
    import torch
    n, batch = 1024, 5
    x = crandn((batch, n), 3)
    ref = oracle.c2c(x)
    p = aclfft.Plan.plan1d(n, aclfft.Z2Z, batch)
    y = np.zeros_like(x)
    p.exec_z2z(x, y, aclfft.FORWARD)  # host memory: staged, synchronous
    assert oracle.rel_l2(y, ref) <= fp64_tol(n)
    xt = torch.from_numpy(x).cuda()
    p.exec_z2z(xt, xt, aclfft.FORWARD)  # in place on the device
    torch.cuda.synchronize()
    assert oracle.rel_l2(xt.cpu().numpy(), ref) <= fp64_tol(n)
    p.destroy()


@pytest.mark.parametrize("n", [4096, 243])
def test_double_real_adjoints(cuda, n):
    """<D2Z x, g> = <x, D2Z^T g> with the real inner product over the half spectrum: the adjoint
    entry points on double plans, checked against the identity in FP64."""
    import torch
    batch = 3
    h = n // 2 + 1
    rng = np.random.default_rng(5)
    x = rng.standard_normal((batch, n))
    g = crandn((batch, h), 6)
    p = aclfft.Plan.plan1d(n, aclfft.D2Z, batch)
    X = torch.empty((batch, h), dtype=torch.complex128, device="cuda")
    p.exec_d2z(torch.from_numpy(x).cuda(), X)
    gx = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    p.exec_r2c_adjoint(torch.from_numpy(g).cuda(), gx)
    torch.cuda.synchronize()
    lhs = np.sum((X.cpu().numpy().conj() * g).real)
    rhs = np.sum(x * gx.cpu().numpy())
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.linalg.norm(x) * np.linalg.norm(g))
    p.destroy()


def _z_variants():
    seen, out = {}, []
    for k in aclfft.kernel_registry():
        if k["kind"] not in (8, 9, 10, 11, 12):
            continue
        key = (k["kind"], k["n"])
        idx = seen.get(key, 0)
        seen[key] = idx + 1
        out.append((k["kind"], k["n"], idx))
    return out


@pytest.mark.parametrize("kind,n,variant", _z_variants())
def test_every_double_kernel(cuda, monkeypatch, kind, n, variant):
    """Each compiled complex128 configuration, forced with ACLFFT_KERNEL_VARIANT."""
    import torch
    monkeypatch.setenv("ACLFFT_KERNEL_VARIANT", str(variant))
    batch = 11
    if kind == 8:
        x = crandn((batch, n), n + variant)
        p = aclfft.Plan.plan1d(n, aclfft.Z2Z, batch)
        y = torch.empty((batch, n), dtype=torch.complex128, device="cuda")
        p.exec_z2z(torch.from_numpy(x).cuda(), y, aclfft.BACKWARD)
        torch.cuda.synchronize()
        err = oracle.rel_l2(y.cpu().numpy(), oracle.c2c(x, aclfft.BACKWARD))
        m = n
    elif kind in (9, 11):
        m = 2 * n if kind == 9 else n
        x = np.random.default_rng(variant).standard_normal((batch, m))
        p = aclfft.Plan.plan1d(m, aclfft.D2Z, batch)
        y = torch.empty((batch, m // 2 + 1), dtype=torch.complex128, device="cuda")
        p.exec_d2z(torch.from_numpy(x).cuda(), y)
        torch.cuda.synchronize()
        err = oracle.rel_l2(y.cpu().numpy(), oracle.r2c(x))
    else:
        m = 2 * n if kind == 10 else n
        X = crandn((batch, m // 2 + 1), variant)
        X[:, 0] = X[:, 0].real
        if m % 2 == 0:
            X[:, -1] = X[:, -1].real
        p = aclfft.Plan.plan1d(m, aclfft.Z2D, batch)
        y = torch.empty((batch, m), dtype=torch.float64, device="cuda")
        p.exec_z2d(torch.from_numpy(X).cuda(), y)
        torch.cuda.synchronize()
        err = oracle.rel_l2(y.cpu().numpy(), oracle.c2r(X, m))
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    assert err <= fp64_tol(m), err
