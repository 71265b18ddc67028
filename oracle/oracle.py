"""Python access to the FP64 oracle (fft_oracle.c) and the reference's verdict semantics.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, never by the product path (ops-fft_amd/).

Array conventions: complex data as numpy complex128 arrays (interleaved re/im in memory), real
data as float64.  Transforms are unnormalised in both directions, like the aclfft API.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def load(path):
    """ctypes handle of an oracle build at `path` with the argument types set."""
    L = ctypes.CDLL(path)
    d = ctypes.POINTER(ctypes.c_double)
    ll = ctypes.c_longlong
    for name, args in (
        ("oracle_dft_naive", [d, d, ll, ll, ctypes.c_int]),
        ("oracle_c2c", [d, d, ll, ll, ctypes.c_int]),
        ("oracle_r2c", [d, d, ll, ll]),
        ("oracle_c2r", [d, d, ll, ll]),
        ("oracle_c2c_2d", [d, d, ll, ll, ll, ctypes.c_int]),
        ("oracle_c2c_vertical", [d, d, ll, ll, ctypes.c_int]),
        ("oracle_c2c_3d", [d, d, ll, ll, ll, ll, ctypes.c_int]),
    ):
        f = getattr(L, name)
        f.argtypes = args
        f.restype = None
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = load(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c128(x):
    return np.ascontiguousarray(x, dtype=np.complex128)


FORWARD = -1
BACKWARD = 1


def c2c(x, direction=FORWARD):
    """Batched 1D C2C over the last axis."""
    x = _c128(x)
    n = x.shape[-1]
    out = np.empty_like(x)
    lib().oracle_c2c(_p(x), _p(out), n, x.size // max(n, 1), direction)
    return out


def dft_naive(x, direction=FORWARD):
    x = _c128(x)
    n = x.shape[-1]
    out = np.empty_like(x)
    lib().oracle_dft_naive(_p(x), _p(out), n, x.size // max(n, 1), direction)
    return out


def r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[-1]
    out = np.empty(x.shape[:-1] + (n // 2 + 1,), dtype=np.complex128)
    lib().oracle_r2c(_p(x), _p(out), n, x.size // n)
    return out


def c2r(X, n):
    X = _c128(X)
    assert X.shape[-1] == n // 2 + 1
    out = np.empty(X.shape[:-1] + (n,), dtype=np.float64)
    lib().oracle_c2r(_p(X), _p(out), n, X.size // (n // 2 + 1))
    return out


def c2c_2d(x, direction=FORWARD):
    """2D C2C over the last two axes ([..., nx, ny], ny fastest)."""
    x = _c128(x)
    nx, ny = x.shape[-2], x.shape[-1]
    out = np.empty_like(x)
    lib().oracle_c2c_2d(_p(x), _p(out), nx, ny, x.size // (nx * ny), direction)
    return out


def c2c_3d(x, direction=FORWARD):
    """3D C2C over the last three axes ([..., nx, ny, nz], nz fastest)."""
    x = _c128(x)
    nx, ny, nz = x.shape[-3], x.shape[-2], x.shape[-1]
    out = np.empty_like(x)
    lib().oracle_c2c_3d(_p(x), _p(out), nx, ny, nz, x.size // (nx * ny * nz), direction)
    return out


def c2c_vertical(x, direction=FORWARD):
    """Column transforms of an [n, cols] matrix (the aclfft VERTICAL layout)."""
    x = _c128(x)
    n, cols = x.shape
    out = np.empty_like(x)
    lib().oracle_c2c_vertical(_p(x), _p(out), n, cols, direction)
    return out


# ------------------------------------------------------------------------------------------------
# verdict semantics of the reference harness (src/common/host/test_helper.py:41-71, 28-38, 97-103):
# computed in f64 on the flattened float32 views; empty data, size mismatch or non-finite values
# fail; then any |o - g| > atol fails; then rel-L2 ||o - g|| / ||g|| > rtol fails (absolute L2
# when ||g|| = 0).
# ------------------------------------------------------------------------------------------------
def flat_f32_view(a):
    a = np.asarray(a)
    if np.iscomplexobj(a):
        a = np.ascontiguousarray(a).view(a.real.dtype)
    return np.asarray(a, dtype=np.float64).reshape(-1)


def comparison_details(output, golden, atol, rtol):
    o = flat_f32_view(output)
    g = flat_f32_view(golden)
    if g.size == 0 or o.size == 0:
        return False, "empty comparison data (golden=%d, output=%d)" % (g.size, o.size)
    if g.size != o.size:
        return False, "size mismatch (golden=%d, output=%d)" % (g.size, o.size)
    if not (np.all(np.isfinite(o)) and np.all(np.isfinite(g))):
        return False, "nonfinite comparison values"
    diff = np.abs(o - g)
    max_abs = float(diff.max())
    gn = float(np.linalg.norm(g))
    rel = float(np.linalg.norm(o - g) / gn) if gn > 0 else float(np.linalg.norm(o - g))
    bad = int(np.sum(diff > atol))
    if bad > 0:
        i = int(np.nonzero(diff > atol)[0][0])
        return False, "%d/%d exceed atol=%g, max_abs_err=%.6e, rel_l2=%.6e, first at idx=%d" % (
            bad, g.size, atol, max_abs, rel, i)
    if rel > rtol:
        return False, "rel_l2=%.6e > rtol=%g, max_abs_err=%.6e" % (rel, rtol, max_abs)
    return True, "max_abs_err=%.6e, rel_l2=%.6e (atol=%g, rtol=%g)" % (max_abs, rel, atol, rtol)


def rel_l2(output, reference):
    o = np.asarray(output, dtype=np.complex128 if np.iscomplexobj(output) else np.float64)
    r = np.asarray(reference, dtype=o.dtype)
    rn = np.linalg.norm(r.reshape(-1))
    d = np.linalg.norm((o - r).reshape(-1))
    return float(d / rn) if rn > 0 else float(d)
