"""Shared helpers for the tests (seeded inputs, tolerances)."""
import math

import numpy as np


def crandn(shape, seed):
    """complex128 with re, im ~ N(0, 1/2) (the distribution of torch.randn(complex64))."""
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * math.sqrt(0.5)


def fp32_tol(n):
    """The FP32 path's bar against the FP64 oracle: rel-L2 <= 1e-6 * log2(N)."""
    return 1e-6 * max(1.0, math.log2(max(n, 2)))


def col_variants(kind=3):
    """(rows, variant index) of every compiled column kernel of `kind` (3: complex64,
    14: complex128); the index counts entries of the same kind and length, as
    ACLFFT_COLUMN_VARIANT does."""
    import ctypes

    import aclfft
    lib = aclfft.lib
    lib.ofx_col_kernel_count.restype = ctypes.c_int
    out, seen = [], {}
    for i in range(lib.ofx_col_kernel_count()):
        info = aclfft._lib.KernelInfo()
        lib.ofx_col_kernel_info_get(i, ctypes.byref(info))
        if info.kind != kind:
            continue
        idx = seen.get(info.n, 0)
        seen[info.n] = idx + 1
        out.append((info.n, idx))
    return out
