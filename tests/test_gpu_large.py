"""GPU parity of the large-length C2C paths (four-step: column kernel with fused W_N twiddles, then
row kernel with transposed store; or the runtime-radix kernel for lengths without compiled
factors) against the FP64 oracle, both directions, batch > 1 so that images do not overlap."""
import numpy as np
import pytest

import aclfft
from oracle import oracle
from _util import crandn, fp32_tol

pytestmark = pytest.mark.gpu

LARGE = [
    (1 << 14, 3), (1 << 15, 3), (1 << 16, 2), (1 << 17, 2), (1 << 18, 2), (1 << 20, 2),
    (1 << 21, 2), (1 << 22, 1), (1 << 24, 1), (3 << 14, 2), (1000000, 1), (46200, 3),
]


@pytest.mark.parametrize("n,batch", LARGE)
@pytest.mark.parametrize("direction", [aclfft.FORWARD, aclfft.BACKWARD])
def test_large_c2c(cuda, n, batch, direction):
    import torch
    x = crandn((batch, n), seed=n % 1000003 + direction).astype(np.complex64)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p.exec_c2c(xt, yt, direction)
    torch.cuda.synchronize()
    desc = p.describe()
    p.destroy()
    ref = oracle.c2c(x.astype(np.complex128), direction)
    err = oracle.rel_l2(yt.cpu().numpy(), ref)
    print("n=%d batch=%d dir=%d rel_l2=%.3e\n%s" % (n, batch, direction, err, desc))
    assert err <= fp32_tol(n), err


@pytest.mark.parametrize("n", [1 << 16, 1 << 20])
def test_large_c2c_in_place(cuda, n):
    """The four-step path with out == in (the row step writes the caller's buffer from the
    plan's workspace)."""
    import torch
    batch = 2
    x = crandn((batch, n), seed=7).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    p.exec_c2c(xt, xt, aclfft.FORWARD)
    torch.cuda.synchronize()
    p.destroy()
    err = oracle.rel_l2(xt.cpu().numpy(), oracle.c2c(x.astype(np.complex128)))
    assert err <= fp32_tol(n), err


@pytest.mark.parametrize("n,batch,levels", [(1 << 19, 3, 3), (1 << 20, 2, 3), (1 << 22, 1, 2), (1 << 25, 1, 0)])
def test_fourstep_forms(cuda, monkeypatch, n, batch, levels):
    """The two- and three-factor four-step forms forced where both exist (ACLFFT_FOURSTEP_LEVELS),
    with batch > 1 for the three-factor form's [k2][b][k1][n3] workspace layout."""
    import torch
    monkeypatch.setenv("ACLFFT_FOURSTEP_LEVELS", str(levels))
    x = crandn((batch, n), seed=levels + n % 97).astype(np.complex64)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    desc = p.describe()
    if levels:
        assert desc.count("fourstep-twiddle") == levels - 1, desc
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    p.exec_c2c(xt, yt, aclfft.BACKWARD)
    torch.cuda.synchronize()
    p.destroy()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c(x.astype(np.complex128), aclfft.BACKWARD))
    print(desc, "rel_l2 %.3e" % err)
    assert err <= fp32_tol(n), err


def test_max_length_properties(cuda):
    """n = 2^27, the largest horizontal length a plan accepts.  The oracle takes minutes here, so
    the check is size-independent: Parseval, the forward/backward round trip (backward is
    unnormalised: x -> n x), and a few bins against their defining sums in FP64."""
    import torch
    n = 1 << 27
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(n, dtype=torch.complex64, device="cuda", generator=g)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, 1)
    X = torch.empty_like(x)
    p.exec_c2c(x, X, aclfft.FORWARD)
    torch.cuda.synchronize()
    ex = float((x.abs().double() ** 2).sum())
    eX = float((X.abs().double() ** 2).sum())
    assert abs(eX / (n * ex) - 1.0) < 1e-5
    y = torch.empty_like(x)
    p.exec_c2c(X, y, aclfft.BACKWARD)
    torch.cuda.synchronize()
    rt = float(torch.linalg.vector_norm((y / n - x).to(torch.complex128)) / torch.linalg.vector_norm(x.to(torch.complex128)))
    print("2^27 round trip rel_l2 %.3e" % rt)
    assert rt <= 2 * fp32_tol(n), rt
    # spot bins: X[k] = sum_j x[j] exp(-2 pi i j k / n), exact integer phase (j*k mod n)
    xd = x.to(torch.complex128)
    j = torch.arange(n, dtype=torch.int64, device="cuda")
    for k in (0, 1, 12345, n // 2 + 7, n - 1):
        ph = (j * k) % n
        w = torch.polar(torch.ones(n, dtype=torch.float64, device="cuda"), ph.double() * (-2.0 * np.pi / n))
        ref = complex((xd * w).sum())
        got = complex(X[k].to(torch.complex128).item())
        err = abs(got - ref) / np.sqrt(n)
        print("bin %d err/sqrt(n) %.3e" % (k, err))
        assert err < 1e-4, (k, got, ref)
    p.destroy()


def _rowt_variants():
    seen, out = {}, []
    for k in aclfft.kernel_registry():
        if k["kind"] not in (6, 13):  # transposed-store row kernels, complex64 and complex128
            continue
        key = (k["kind"], k["n"])
        out.append((k["kind"], k["n"], seen.get(key, 0)))
        seen[key] = seen.get(key, 0) + 1
    return out


@pytest.mark.parametrize("kind,n2,variant", _rowt_variants())
def test_every_rowt_kernel(cuda, monkeypatch, kind, n2, variant):
    """Each compiled transposed-store row kernel as the row step of a two-factor four-step
    (ACLFFT_FOURSTEP_LEVELS=2, row length n2), forced with ACLFFT_KERNEL_VARIANT."""
    import torch
    z = kind == 13
    monkeypatch.setenv("ACLFFT_FOURSTEP_LEVELS", "2")
    monkeypatch.setenv("ACLFFT_KERNEL_VARIANT", str(variant))
    n = n2 * n2  # the balanced two-factor split: columns and rows of n2 points
    batch = 2
    x = crandn((batch, n), n2 + variant)
    if not z:
        x = x.astype(np.complex64)
    p = aclfft.Plan.plan1d(n, aclfft.Z2Z if z else aclfft.C2C, batch)
    desc = p.describe()
    if ("kind13" if z else "kind6") + " n=%d " % n2 not in desc:
        p.destroy()
        pytest.skip("the four-step of %d does not end in a %d-point row step: %s" % (n, n2, desc))
    xt = torch.from_numpy(x).cuda()
    yt = torch.empty_like(xt)
    (p.exec_z2z if z else p.exec_c2c)(xt, yt, aclfft.BACKWARD)
    torch.cuda.synchronize()
    err = oracle.rel_l2(yt.cpu().numpy(), oracle.c2c(x.astype(np.complex128), aclfft.BACKWARD))
    print(desc, "rel_l2 %.3e" % err)
    p.destroy()
    tol = 1e-14 * np.log2(n) if z else fp32_tol(n)
    assert err <= tol, err


@pytest.mark.parametrize("n,batch", [(1 << 20, 2), (1 << 21, 2), (1 << 22, 1), (1 << 24, 1), (1 << 19, 3)])
def test_fourstep_twiddle_precision(cuda, monkeypatch, n, batch):
    """Four-step twiddles: formed in FP64 and rounded once (the default for n >= 2^20) or from
    FP32 factors (about 2 ulp; ACLFFT_FOURSTEP_FP64=0); both meet the FP32 bar against the FP64
    oracle, and the FP64 form is no worse than the FP32 form beyond noise."""
    import torch
    x = crandn((batch, n), seed=n % 7919).astype(np.complex64)
    ref = oracle.c2c(x.astype(np.complex128))
    xt = torch.from_numpy(x).cuda()
    errs = {}
    for f in ("0", "1"):
        monkeypatch.setenv("ACLFFT_FOURSTEP_FP64", f)
        p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
        desc = p.describe()
        assert ("fourstep-twiddle-fp64" in desc) == (f == "1"), desc
        yt = torch.empty_like(xt)
        p.exec_c2c(xt, yt, aclfft.FORWARD)
        torch.cuda.synchronize()
        p.destroy()
        errs[f] = oracle.rel_l2(yt.cpu().numpy(), ref)
        assert errs[f] <= fp32_tol(n), (f, errs[f])
    monkeypatch.delenv("ACLFFT_FOURSTEP_FP64")
    p = aclfft.Plan.plan1d(n, aclfft.C2C, batch)
    assert ("fourstep-twiddle-fp64" in p.describe()) == (n >= (1 << 20))
    p.destroy()
    print("n=%d: rel-L2 FP32 factored twiddles %.4e, FP64 twiddles %.4e" % (n, errs["0"], errs["1"]))
    assert errs["1"] <= errs["0"] * 1.05
