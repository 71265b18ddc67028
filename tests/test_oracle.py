"""Pins of the FP64 oracle (CPU only).

The oracle restates the reference's algorithms (oracle/fft_oracle.c).  It is checked here against
what the mathematics fixes and against the reference's golden path, so that a plausible mistake
(a dropped term, a wrong sign or index, a transposed operand) fails at least one test:
  - brute-force long-double DFT (the definition) for every n <= 64 and mixed sizes up to 1000,
  - closed forms: delta, constant, single tone (catches the exponent sign), shifted delta (index),
  - linearity, Parseval, the unnormalised round trip backward(forward(x)) = n x,
  - numpy.fft (pocketfft, a library routine) at large and prime-power sizes,
  - R2C = truncated C2C of real input, exact zero DC/Nyquist imaginary parts, C2R(R2C(x)) = n x,
    the Nyquist-only spectrum -> (-1)^m (the reference's fixture, irfft1_d gen_data.py:54-56),
  - 2D separability and the VERTICAL layout as a transposed batch,
  - the committed torch.fft goldens of the reference's own cases (tests/golden/), with the
    reference's verdict and a tight bound,
  - a build of the oracle with a planted twiddle-sign mistake fails the pins.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle
import ref_cases as rc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_goldens.npz")


def crand(shape, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


@pytest.mark.parametrize("n", list(range(1, 65)) + [81, 96, 105, 121, 143, 210, 243, 420, 1000])
def test_c2c_matches_brute_force(n):
    x = crand((2, n), n)
    for d in (oracle.FORWARD, oracle.BACKWARD):
        err = oracle.rel_l2(oracle.c2c(x, d), oracle.dft_naive(x, d))
        assert err < 1e-13, (n, d, err)


def test_brute_force_definition_small():
    # the definition itself, evaluated independently in numpy for n = 5
    n = 5
    x = crand((1, n), 1)
    j = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(j, j) / n)
    assert np.allclose(oracle.dft_naive(x)[0], F @ x[0], atol=1e-14)
    assert np.allclose(oracle.dft_naive(x, oracle.BACKWARD)[0], F.conj() @ x[0], atol=1e-14)


@pytest.mark.parametrize("n", [1, 2, 7, 16, 60, 1000, 1024, 3000])
def test_closed_forms(n):
    k = np.arange(n)
    delta = np.zeros(n, complex)
    delta[0] = 1
    assert np.allclose(oracle.c2c(delta), np.ones(n), atol=1e-13)
    assert np.allclose(oracle.c2c(np.ones(n, complex)), n * delta, atol=1e-9 * n)
    f = 3 % n
    tone = np.exp(2j * np.pi * f * k / n)  # forward transform: n at bin f (sign of the exponent)
    expect = np.zeros(n, complex)
    expect[f] = n
    assert np.allclose(oracle.c2c(tone), expect, atol=1e-9 * n)
    s = 5 % n
    shifted = np.zeros(n, complex)
    shifted[s] = 1
    assert np.allclose(oracle.c2c(shifted), np.exp(-2j * np.pi * k * s / n), atol=1e-12)


@pytest.mark.parametrize("n", [12, 243, 1024, 2310])
def test_linearity_parseval_roundtrip(n):
    x, y = crand((3, n), 1), crand((3, n), 2)
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    X, Y = oracle.c2c(x), oracle.c2c(y)
    assert oracle.rel_l2(oracle.c2c(a * x + b * y), a * X + b * Y) < 1e-13
    assert np.allclose(np.sum(np.abs(X) ** 2, -1), n * np.sum(np.abs(x) ** 2, -1), rtol=1e-12)
    assert oracle.rel_l2(oracle.c2c(X, oracle.BACKWARD), n * x) < 1e-13


@pytest.mark.parametrize("n", [4096, 3 ** 8, 5 ** 5, 2 * 3 * 5 * 7 * 11 * 13, 46200, 97, 1 << 16])
def test_c2c_matches_numpy(n):
    x = crand((2, n), n)
    assert oracle.rel_l2(oracle.c2c(x), np.fft.fft(x)) < 1e-13
    assert oracle.rel_l2(oracle.c2c(x, oracle.BACKWARD), np.fft.ifft(x) * n) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 3, 7, 8, 15, 16, 255, 999, 1024, 1025, 3000, 4096, 8190])
def test_r2c_c2r(n):
    r = np.random.default_rng(n)
    x = r.standard_normal((3, n))
    X = oracle.r2c(x)
    full = oracle.c2c(x.astype(complex))
    assert oracle.rel_l2(X, full[:, : n // 2 + 1]) < 1e-13
    assert np.all(X[:, 0].imag == 0)
    if n % 2 == 0:
        assert np.all(X[:, -1].imag == 0)
    assert oracle.rel_l2(X, np.fft.rfft(x)) < 1e-13
    assert oracle.rel_l2(oracle.c2r(X, n), n * x) < 1e-13
    # imaginary parts of DC (and Nyquist) are ignored by C2R
    Xd = X.copy()
    Xd[:, 0] += 5j
    if n % 2 == 0:
        Xd[:, -1] -= 3j
    assert oracle.rel_l2(oracle.c2r(Xd, n), n * x) < 1e-13


def test_c2r_nyquist_only():
    n = 8192
    X = np.zeros((1, n // 2 + 1), complex)
    X[0, -1] = 1
    assert np.array_equal(oracle.c2r(X, n)[0], np.where(np.arange(n) % 2 == 0, 1.0, -1.0))


def test_2d_separable_and_vertical():
    nx, ny = 12, 20
    a, b = crand(nx, 1), crand(ny, 2)
    img = np.outer(a, b)[None]
    expect = np.outer(oracle.c2c(a), oracle.c2c(b))[None]
    assert oracle.rel_l2(oracle.c2c_2d(img), expect) < 1e-13
    x = crand((2, nx, ny), 3)
    assert oracle.rel_l2(oracle.c2c_2d(x), np.fft.fft2(x)) < 1e-13
    assert oracle.rel_l2(oracle.c2c_2d(x, oracle.BACKWARD), np.fft.ifft2(x) * nx * ny) < 1e-13
    m = crand((64, 7), 4)
    assert oracle.rel_l2(oracle.c2c_vertical(m), oracle.c2c(m.T).T) < 1e-14


def test_3d_separable_delta_and_fftn():
    """3D: separable inputs transform to the outer product of 1D transforms; a delta at (i, j, k)
    to the closed-form phase ramp; random data matches numpy.fft.fftn / ifftn * N over the last
    three axes (distinct nx, ny, nz, so a swapped axis fails)."""
    nx, ny, nz = 6, 10, 12
    a, b, c = crand(nx, 5), crand(ny, 6), crand(nz, 7)
    vol = np.einsum("i,j,k->ijk", a, b, c)[None]
    expect = np.einsum("i,j,k->ijk", oracle.c2c(a), oracle.c2c(b), oracle.c2c(c))[None]
    assert oracle.rel_l2(oracle.c2c_3d(vol), expect) < 1e-13
    d = np.zeros((1, nx, ny, nz), complex)
    d[0, 2, 3, 5] = 1
    u, v, w = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ramp = np.exp(-2j * np.pi * (2 * u / nx + 3 * v / ny + 5 * w / nz))
    assert np.max(np.abs(oracle.c2c_3d(d)[0] - ramp)) < 1e-13
    x = crand((2, nx, ny, nz), 8)
    assert oracle.rel_l2(oracle.c2c_3d(x), np.fft.fftn(x, axes=(1, 2, 3))) < 1e-13
    assert oracle.rel_l2(oracle.c2c_3d(x, oracle.BACKWARD),
                         np.fft.ifftn(x, axes=(1, 2, 3)) * nx * ny * nz) < 1e-13


# ---- the reference's golden path (torch.fft), committed fixtures -------------------------------
def _golden():
    return np.load(GOLDEN, allow_pickle=False)


def test_fixture_meta():
    g = _golden()
    assert "seed 12345" in str(g["__meta__"])


@pytest.mark.parametrize("case", rc.C2C, ids=[c[0] for c in rc.C2C])
def test_oracle_vs_reference_goldens_c2c(case):
    name, n, batch, mode, atol, rtol = case
    g = _golden()
    key = "c2c/%s/" % name
    if key + "in" not in g.files:
        pytest.skip("not in the committed fixtures (size)")
    x = g[key + "in"].astype(np.complex128)
    if mode == "vertical":
        out = oracle.c2c_vertical(x.reshape(n, -1)).reshape(x.shape)
    elif mode == "roundtrip":
        out = oracle.c2c(oracle.c2c(x), oracle.BACKWARD)
    else:
        out = oracle.c2c(x, oracle.FORWARD if mode == "forward" else oracle.BACKWARD)
    ok, detail = oracle.comparison_details(out, g[key + "golden"], atol, rtol)
    assert ok, detail
    assert oracle.rel_l2(out, g[key + "golden"]) < 1e-6  # golden is complex64


@pytest.mark.parametrize("case", rc.R2C, ids=[c[0] for c in rc.R2C])
def test_oracle_vs_reference_goldens_r2c(case):
    name, n, batch, atol, rtol = case
    g = _golden()
    key = "r2c/%s/" % name
    if key + "in" not in g.files:
        pytest.skip("not in the committed fixtures (size)")
    out = oracle.r2c(g[key + "in"].astype(np.float64))
    ok, detail = oracle.comparison_details(out, g[key + "golden"], atol, rtol)
    assert ok, detail
    assert oracle.rel_l2(out, g[key + "golden"]) < 1e-6


@pytest.mark.parametrize("case", rc.C2R, ids=[c[0] for c in rc.C2R])
def test_oracle_vs_reference_goldens_c2r(case):
    name, n, batch, nyq, atol, rtol = case
    g = _golden()
    key = "c2r/%s/" % name
    if key + "in" not in g.files:
        pytest.skip("not in the committed fixtures (size)")
    out = oracle.c2r(g[key + "in"].astype(np.complex128), n)
    ok, detail = oracle.comparison_details(out, g[key + "golden"], atol, rtol)
    assert ok, detail
    assert oracle.rel_l2(out, g[key + "golden"]) < 1e-6


@pytest.mark.parametrize("case", rc.FFT2, ids=[c[0] for c in rc.FFT2])
def test_oracle_vs_reference_goldens_fft2(case):
    name, nx, ny, batch, fwd, atol, rtol = case
    g = _golden()
    key = "fft2/%s/" % name
    if key + "in" not in g.files:
        pytest.skip("not in the committed fixtures (size)")
    out = oracle.c2c_2d(g[key + "in"].astype(np.complex128), oracle.FORWARD if fwd else oracle.BACKWARD)
    ok, detail = oracle.comparison_details(out, g[key + "golden"], atol, rtol)
    assert ok, detail
    assert oracle.rel_l2(out, g[key + "golden"]) < 1e-6


# ---- verdict semantics (reference harness tests: tests/host/test_fft_nan_tolerance_cli.py,
#      tests/python/test_comparison_artifacts.py) ----------------------------------------------
def test_verdict_semantics():
    g = np.array([1.0, 2.0, 3.0], np.float32)
    assert oracle.comparison_details(g, g, 0.0, 0.0)[0]
    assert not oracle.comparison_details(np.array([], np.float32), g, 1, 1)[0]       # empty
    assert not oracle.comparison_details(g[:2], g, 1, 1)[0]                           # size
    assert not oracle.comparison_details(np.array([1, np.nan, 3], np.float32), g, 1, 1)[0]
    assert not oracle.comparison_details(np.array([1, np.inf, 3], np.float32), g, 1e9, 1e9)[0]
    o = g + np.array([0, 0, 0.5], np.float32)
    ok, d = oracle.comparison_details(o, g, 0.4, 1.0)
    assert not ok and "exceed atol" in d                                              # atol first
    ok, d = oracle.comparison_details(o, g, 0.6, 0.1)
    assert not ok and "rtol" in d                                                     # then rel-L2
    assert oracle.comparison_details(o, g, 0.6, 0.2)[0]
    z = np.zeros(3, np.float32)
    assert not oracle.comparison_details(z + 0.1, z, 1.0, 0.1)[0]  # |g| = 0: absolute L2 vs rtol


# ---- the pins have teeth -----------------------------------------------------------------------
def test_planted_twiddle_sign_mistake_is_caught(tmp_path):
    so = tmp_path / "liboracle_planted.so"
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-DOR_PLANT_TWIDDLE_SIGN", "-o", str(so),
                           os.path.join(ROOT, "oracle", "fft_oracle.c"), "-lm"])
    bad = oracle.load(str(so))
    n = 12
    x = np.ascontiguousarray(crand((1, n), 0))
    out = np.empty_like(x)
    bad.oracle_c2c(oracle._p(x), oracle._p(out), n, 1, -1)
    assert oracle.rel_l2(out, oracle.dft_naive(x)) > 1e-3  # brute-force pin fails
    tone = np.ascontiguousarray(np.exp(2j * np.pi * 3 * np.arange(n) / n)[None])
    bad.oracle_c2c(oracle._p(tone), oracle._p(out), n, 1, -1)
    expect = np.zeros(n, complex)
    expect[3] = n
    assert not np.allclose(out[0], expect, atol=1e-6)  # closed-form pin fails
