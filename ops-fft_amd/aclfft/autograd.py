"""torch.autograd.Function wrappers over the aclfft C API (ctypes), for CUDA (ROCm) tensors.

Functions follow torch.fft's conventions (last axis or last two axes, norm = "backward" |
"forward" | "ortho") so that a model can swap them in:

    fft(x), ifft(x), rfft(x), irfft(X, n), fft2(x), ifft2(x)

Forward passes run the library's FP32 kernels with the normalisation fused into the last launch
(aclfftExtSetScale).  Backward passes run the hand-written adjoints:
  - C2C: the transform in the opposite direction with the same scale (F^H = conj(F)^T, F symmetric);
  - R2C: aclfftExtExecR2CAdjoint (C2R kernels, bins strictly between DC and Nyquist weighted 1/2);
  - C2R: aclfftExtExecC2RAdjoint (R2C kernels, the same bins weighted 2).
The gradient convention is PyTorch's (grad = dL/dRe + i dL/dIm).  No CPU fallback exists: the
library must be built and the tensors must live on the GPU.
"""
import math
import threading

import torch

from . import _lib
from .plan import Plan

_cache = {}
_lock = threading.Lock()


def _scale(norm, n, inverse):
    if norm is None or norm == "backward":
        return 1.0 / n if inverse else 1.0
    if norm == "forward":
        return 1.0 if inverse else 1.0 / n
    if norm == "ortho":
        return 1.0 / math.sqrt(n)
    raise ValueError("norm must be 'backward', 'forward' or 'ortho'")


def _plan(kind, dims, batch, scale, device):
    key = (kind, dims, batch, scale, device)
    with _lock:
        p = _cache.get(key)
        if p is None:
            if kind == "c2c":
                p = Plan.plan1d(dims[0], _lib.C2C, batch)
            elif kind == "r2c":
                p = Plan.plan1d(dims[0], _lib.R2C, batch)
            elif kind == "c2r":
                p = Plan.plan1d(dims[0], _lib.C2R, batch)
            else:  # 2d
                p = Plan.plan2d(batch, dims[0], dims[1], _lib.C2C)
            if scale != 1.0:
                p.set_scale(scale)
            _cache[key] = p
    return p


def _check(x, dtype):
    if not x.is_cuda:
        raise ValueError("aclfft.autograd works on GPU tensors only")
    if x.dtype != dtype:
        raise TypeError("expected %s, got %s" % (dtype, x.dtype))
    return x.contiguous()


def _c2c(x, direction, scale, two_d):
    x = _check(x, torch.complex64)
    if two_d:
        nx, ny = x.shape[-2], x.shape[-1]
        p = _plan("2d", (nx, ny), x.numel() // (nx * ny), scale, x.device.index)
    else:
        n = x.shape[-1]
        p = _plan("c2c", (n,), x.numel() // n, scale, x.device.index)
    y = torch.empty_like(x)
    p.exec_c2c(x, y, direction)
    return y


class _C2C(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, direction, scale, two_d):
        ctx.direction, ctx.scale, ctx.two_d = direction, scale, two_d
        return _c2c(x, direction, scale, two_d)

    @staticmethod
    def backward(ctx, g):
        return _c2c(g, -ctx.direction, ctx.scale, ctx.two_d), None, None, None


class _R2C(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, scale):
        x = _check(x, torch.float32)
        n = x.shape[-1]
        b = x.numel() // n
        p = _plan("r2c", (n,), b, scale, x.device.index)
        y = torch.empty(x.shape[:-1] + (n // 2 + 1,), dtype=torch.complex64, device=x.device)
        p.exec_r2c(x, y)
        ctx.n, ctx.b, ctx.scale, ctx.shape = n, b, scale, x.shape
        return y

    @staticmethod
    def backward(ctx, g):
        g = _check(g, torch.complex64)
        p = _plan("r2c", (ctx.n,), ctx.b, ctx.scale, g.device.index)
        gx = torch.empty(ctx.shape, dtype=torch.float32, device=g.device)
        p.exec_r2c_adjoint(g, gx)
        return gx, None


class _C2R(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, n, scale):
        X = _check(X, torch.complex64)
        if X.shape[-1] != n // 2 + 1:
            raise ValueError("irfft: last axis must hold n//2+1 bins")
        b = X.numel() // X.shape[-1]
        p = _plan("c2r", (n,), b, scale, X.device.index)
        y = torch.empty(X.shape[:-1] + (n,), dtype=torch.float32, device=X.device)
        p.exec_c2r(X, y)
        ctx.n, ctx.b, ctx.scale, ctx.shape = n, b, scale, X.shape
        return y

    @staticmethod
    def backward(ctx, g):
        g = _check(g, torch.float32)
        p = _plan("c2r", (ctx.n,), ctx.b, ctx.scale, g.device.index)
        gX = torch.empty(ctx.shape, dtype=torch.complex64, device=g.device)
        p.exec_c2r_adjoint(g, gX)
        return gX, None, None


def fft(x, norm="backward"):
    return _C2C.apply(x, _lib.FORWARD, _scale(norm, x.shape[-1], False), False)


def ifft(x, norm="backward"):
    return _C2C.apply(x, _lib.BACKWARD, _scale(norm, x.shape[-1], True), False)


def fft2(x, norm="backward"):
    return _C2C.apply(x, _lib.FORWARD, _scale(norm, x.shape[-1] * x.shape[-2], False), True)


def ifft2(x, norm="backward"):
    return _C2C.apply(x, _lib.BACKWARD, _scale(norm, x.shape[-1] * x.shape[-2], True), True)


def rfft(x, norm="backward"):
    return _R2C.apply(x, _scale(norm, x.shape[-1], False))


def irfft(X, n, norm="backward"):
    return _C2R.apply(X, n, _scale(norm, n, True))


def clear_plan_cache():
    with _lock:
        for p in _cache.values():
            p.destroy()
        _cache.clear()
