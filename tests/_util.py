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
