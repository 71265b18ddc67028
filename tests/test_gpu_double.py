"""GPU parity of the double-precision transforms (aclfftExecZ2Z / D2Z / Z2D) against the FP64
oracle.  Tolerance: rel-L2 <= 1e-14 * log2(n) (FP32 arithmetic would give ~1e-7), both
directions, batches that leave partial workgroups, host and device pointers, in place, adjoints,
the complex128 four-step (two and three factors), 2D Z2Z and the VERTICAL Z2Z layout."""
import math

import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import col_variants, crandn

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


@pytest.mark.parametrize("n,batch,levels", [
    (1 << 14, 3, 0), (1 << 16, 2, 0), (1 << 16, 2, 3), (1 << 20, 2, 2), (1 << 20, 1, 3),
    (1 << 22, 1, 0)])
def test_z2z_fourstep_vs_oracle(cuda, monkeypatch, n, batch, levels):
    """Lengths beyond the single-workgroup kernels: the complex128 four-step over compiled column
    and transposed-store row kernels, both factorisations (ACLFFT_FOURSTEP_LEVELS)."""
    import torch
    if levels:
        monkeypatch.setenv("ACLFFT_FOURSTEP_LEVELS", str(levels))
    x = crandn((batch, n), n + levels)
    p = aclfft.Plan.plan1d(n, aclfft.Z2Z, batch)
    if "not implemented" in p.describe():
        p.destroy()
        pytest.skip("no complex128 factorisation of %d" % n)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    for direction in (aclfft.FORWARD, aclfft.BACKWARD):
        p.exec_z2z(xt, yt, direction)
        torch.cuda.synchronize()
        err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c(x, direction))
        print(p.describe(), "dir", direction, "rel_l2 %.3e" % err)
        assert err <= fp64_tol(n), err
    p.destroy()


def test_z2z_fourstep_in_place(cuda):
    import torch
    n, batch = 1 << 18, 2
    x = crandn((batch, n), 9)
    p = aclfft.Plan.plan1d(n, aclfft.Z2Z, batch)
    xt = torch.from_numpy(x).cuda()
    p.exec_z2z(xt, xt, aclfft.BACKWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(xt.cpu().numpy(), oracle.c2c(x, aclfft.BACKWARD))
    p.destroy()
    assert err <= fp64_tol(n), err


@pytest.mark.parametrize("batch,nx,ny", [(3, 64, 64), (2, 256, 512), (1, 2048, 2048), (2, 4096, 60),
                                         (4, 1, 1024), (2, 512, 1), (1, 128, 1 << 16)])
def test_z2z_2d_vs_oracle(cuda, batch, nx, ny):
    """2D Z2Z: complex128 rows (length ny) then complex128 columns (length nx)."""
    import torch
    x = crandn((batch, nx, ny), nx * 7 + ny)
    p = aclfft.Plan.plan2d(batch, nx, ny, aclfft.Z2Z)
    if "not implemented" in p.describe():
        p.destroy()
        pytest.skip("no complex128 kernels for %dx%d" % (nx, ny))
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    for direction in (aclfft.FORWARD, aclfft.BACKWARD):
        p.exec_z2z(xt, yt, direction)
        torch.cuda.synchronize()
        err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c_2d(x, direction))
        print(p.describe(), "dir", direction, "rel_l2 %.3e" % err)
        assert err <= fp64_tol(nx * ny), err
    p.destroy()


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("cols", [40, 7])
def test_z2z_vertical_vs_oracle(cuda, n, cols):
    """VERTICAL Z2Z ([n][cols] row-major, one transform per column), in place."""
    import torch
    v = crandn((n, cols), n + cols)
    p = aclfft.Plan.plan1d(n, aclfft.Z2Z, cols, aclfft.VERTICAL)
    vt = torch.from_numpy(v).cuda()
    p.exec_z2z(vt, vt, aclfft.FORWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(vt.cpu().numpy(), oracle.c2c_vertical(v))
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    assert err <= fp64_tol(n), err


@pytest.mark.parametrize("rows,variant", col_variants(14))
@pytest.mark.parametrize("cols", [60, 243])
def test_every_double_column_kernel(cuda, monkeypatch, rows, variant, cols):
    """Each compiled complex128 column kernel as the second pass of a 2D Z2Z plan, with column
    counts (row lengths that have complex128 row kernels) that leave a partial column group,
    forced with ACLFFT_COLUMN_VARIANT."""
    import torch
    monkeypatch.setenv("ACLFFT_COLUMN_VARIANT", str(variant))
    x = crandn((3, rows, cols), rows + cols + variant)
    p = aclfft.Plan.plan2d(3, rows, cols, aclfft.Z2Z)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p.exec_z2z(xt, yt, aclfft.BACKWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c_2d(x, aclfft.BACKWARD))
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    assert err <= fp64_tol(rows * cols), err


@pytest.mark.parametrize("n,batch", [(16384, 3), (1 << 16, 2), (1 << 20, 1)])
def test_d2z_z2d_fourstep_vs_oracle(cuda, n, batch):
    """Even real lengths beyond the single-workgroup kernels: the complex128 four-step over the n/2
    packed samples plus the Hermitian unpack (D2Z), the pre-pack plus the four-step (Z2D)."""
    import torch
    h = n // 2 + 1
    x = np.random.default_rng(n + 1).standard_normal((batch, n))
    p = aclfft.Plan.plan1d(n, aclfft.D2Z, batch)
    X = torch.empty((batch, h), dtype=torch.complex128, device="cuda")
    p.exec_d2z(torch.from_numpy(x).cuda(), X)
    torch.cuda.synchronize()
    err = oracle.rel_l2(X.cpu().numpy(), oracle.r2c(x))
    print(p.describe(), "d2z rel_l2 %.3e" % err)
    assert err <= fp64_tol(n), err
    G = crandn((batch, h), n + 2)
    G[:, 0] = G[:, 0].real
    G[:, -1] = G[:, -1].real
    q = aclfft.Plan.plan1d(n, aclfft.Z2D, batch)
    y = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    q.exec_z2d(torch.from_numpy(G).cuda(), y)
    torch.cuda.synchronize()
    err = oracle.rel_l2(y.cpu().numpy(), oracle.c2r(G, n))
    print(q.describe(), "z2d rel_l2 %.3e" % err)
    assert err <= fp64_tol(n), err
    # adjoint identity <D2Z x, g> = <x, D2Z^T g> on the four-step path
    gx = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    p.exec_r2c_adjoint(torch.from_numpy(G).cuda(), gx)
    torch.cuda.synchronize()
    lhs = np.sum((X.cpu().numpy().conj() * G).real)
    rhs = np.sum(x * gx.cpu().numpy())
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.linalg.norm(x) * np.linalg.norm(G))
    p.destroy()
    q.destroy()


@pytest.mark.parametrize("n,batch", [(1155, 5), (4608, 2), (2310, 3), (9000, 2), (12000, 2), (3 << 18, 1),
                                     (1000000, 1)])
def test_z2z_runtime_radix_vs_oracle(cuda, n, batch):
    """Lengths without a compiled complex128 kernel: the runtime-radix kernel in FP64 (n <= 4608
    in one step; beyond, the four-step over runtime-radix steps)."""
    import torch
    x = crandn((batch, n), n + 3)
    p = aclfft.Plan.plan1d(n, aclfft.Z2Z, batch)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    for direction in (aclfft.FORWARD, aclfft.BACKWARD):
        p.exec_z2z(xt, yt, direction)
        torch.cuda.synchronize()
        err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c(x, direction))
        print(p.describe(), "dir", direction, "rel_l2 %.3e" % err)
        assert err <= fp64_tol(n), err
    p.destroy()


@pytest.mark.parametrize("n", [1155, 2310, 4095, 10000, 30000])
def test_d2z_z2d_runtime_radix_vs_oracle(cuda, n):
    """D2Z / Z2D through the FP64 runtime-radix kernel: odd n as a full complex transform, even n
    with the Hermitian unpack / pre-pack fused (four-step of the packed samples beyond 9216)."""
    import torch
    batch = 3
    h = n // 2 + 1
    x = np.random.default_rng(n).standard_normal((batch, n))
    p = aclfft.Plan.plan1d(n, aclfft.D2Z, batch)
    X = torch.empty((batch, h), dtype=torch.complex128, device="cuda")
    p.exec_d2z(torch.from_numpy(x).cuda(), X)
    torch.cuda.synchronize()
    err = oracle.rel_l2(X.cpu().numpy(), oracle.r2c(x))
    print(p.describe(), "d2z rel_l2 %.3e" % err)
    assert err <= fp64_tol(n), err
    G = crandn((batch, h), n + 5)
    G[:, 0] = G[:, 0].real
    if n % 2 == 0:
        G[:, -1] = G[:, -1].real
    q = aclfft.Plan.plan1d(n, aclfft.Z2D, batch)
    y = torch.empty((batch, n), dtype=torch.float64, device="cuda")
    q.exec_z2d(torch.from_numpy(G).cuda(), y)
    torch.cuda.synchronize()
    err = oracle.rel_l2(y.cpu().numpy(), oracle.c2r(G, n))
    print(q.describe(), "z2d rel_l2 %.3e" % err)
    p.destroy()
    q.destroy()
    assert err <= fp64_tol(n), err


def test_z2z_2d_3d_runtime_radix(cuda):
    import torch
    x = crandn((2, 100, 60), 11)
    p = aclfft.Plan.plan2d(2, 100, 60, aclfft.Z2Z)
    y = torch.empty((2, 100, 60), dtype=torch.complex128, device="cuda")
    p.exec_z2z(torch.from_numpy(x).cuda(), y, aclfft.BACKWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(y.cpu().numpy(), oracle.c2c_2d(x, aclfft.BACKWARD))
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    assert err <= fp64_tol(6000), err
    v = crandn((30, 20, 10), 12)
    p = aclfft.Plan.plan3d(30, 20, 10, aclfft.Z2Z)
    y = torch.empty((30, 20, 10), dtype=torch.complex128, device="cuda")
    p.exec_z2z(torch.from_numpy(v).cuda(), y, aclfft.FORWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(y.cpu().numpy(), oracle.c2c_3d(v[None])[0])
    print(p.describe(), "rel_l2 %.3e" % err)
    p.destroy()
    assert err <= fp64_tol(6000), err
