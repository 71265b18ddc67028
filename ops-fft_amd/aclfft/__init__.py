"""aclfft for MI355X (gfx950): Python access to libcann_ops_fft.so.

    from aclfft import Plan, C2C, FORWARD
    p = Plan.plan1d(1024, C2C, batch=64)
    p.exec_c2c(x, y, FORWARD)     # x, y: complex64 torch tensors on the GPU or numpy arrays

The plan/exec semantics are those of the aclfft C API (include/cann_ops_fft.h).  The native
library is required; there is no CPU fallback.
"""
from ._lib import (LIB_PATH, AclfftError, check, error_string, kernel_registry, lib, ACLFFT_SYMBOLS,
                   C2C, R2C, C2R, Z2Z, D2Z, Z2D, FORWARD, BACKWARD, HORIZONTAL, VERTICAL,
                   SUCCESS, INVALID_PLAN, ALLOC_FAILED, INVALID_TYPE, INVALID_VALUE, INTERNAL_ERROR,
                   EXEC_FAILED, SETUP_FAILED, INVALID_SIZE, NOT_IMPLEMENTED, NOT_SUPPORTED, INVALID_DEVICE,
                   NO_WORKSPACE, UNALIGNED_DATA,
                   PRECISION_FP32, PRECISION_BF16, PRECISION_FP16X2, PRECISION_BF16X3, PRECISION_AUTO)
from .plan import Plan

__all__ = ["Plan", "AclfftError", "check", "error_string", "kernel_registry", "lib", "LIB_PATH",
           "ACLFFT_SYMBOLS", "C2C", "R2C", "C2R", "Z2Z", "D2Z", "Z2D", "FORWARD", "BACKWARD",
           "HORIZONTAL", "VERTICAL"]
