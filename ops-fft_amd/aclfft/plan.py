"""Python mirror of the aclfft plan interface.

`Plan` wraps an aclfftHandle and exposes the reference's operations with the reference's names and
argument meaning (plan1d(nx, type, batch, dimType), plan2d(batch, nx, ny, type), exec_c2c(idata,
odata, direction), exec_r2c, exec_c2r, set_stream, destroy).  Errors raise AclfftError carrying
the aclfftResult code.

Buffers: torch tensors on a GPU (passed as device pointers; the call is asynchronous on the
plan's stream, which defaults to torch's current stream) or host numpy arrays (the library
stages them and returns after copying the result back, the reference's synchronous semantics).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import lib, check, C2C, R2C, C2R, Z2Z, D2Z, Z2D, FORWARD, BACKWARD, HORIZONTAL, VERTICAL


def _ptr(a):
    """Device or host address of a torch tensor / numpy array (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    # torch.Tensor
    if not a.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return a.data_ptr()


def _is_device(a):
    return a is not None and not isinstance(a, np.ndarray) and a.is_cuda


def _nbytes(a):
    if isinstance(a, np.ndarray):
        return a.nbytes
    return a.numel() * a.element_size()


def _check_size(a, need, what):
    """The C API takes raw pointers and trusts their extent; the Python layer knows the buffer
    sizes and refuses one smaller than the plan reads or writes (a kernel would run past it)."""
    if a is not None and _nbytes(a) < need:
        raise ValueError("%s holds %d bytes, the plan needs %d" % (what, _nbytes(a), need))


class Plan:
    def __init__(self, handle, kind, dims, batch, vertical=False):
        self._h = handle
        self.type = kind
        self.dims = dims
        self.batch = batch
        self.vertical = vertical
        self.work_size = 0
        self._stream = None

    # ---- construction (reference: aclfftPlan1d / aclfftPlan2d / Create + MakePlan*) ----------
    @classmethod
    def plan1d(cls, nx, type=C2C, batch=1, dim_type=HORIZONTAL):
        h = ctypes.c_void_p()
        check(lib.aclfftPlan1d(ctypes.byref(h), nx, type, batch, dim_type), "aclfftPlan1d")
        return cls(h, type, (nx,), batch, dim_type == VERTICAL)

    @classmethod
    def plan2d(cls, batch, nx, ny, type=C2C):
        h = ctypes.c_void_p()
        check(lib.aclfftPlan2d(ctypes.byref(h), batch, nx, ny, type), "aclfftPlan2d")
        return cls(h, type, (nx, ny), batch)

    @classmethod
    def plan3d(cls, nx, ny, nz, type=C2C):
        """One [nx][ny][nz] volume (nz fastest), C2C or Z2Z."""
        h = ctypes.c_void_p()
        check(lib.aclfftPlan3d(ctypes.byref(h), nx, ny, nz, type), "aclfftPlan3d")
        return cls(h, type, (nx, ny, nz), 1)

    @classmethod
    def plan_many_uneven(cls, lengths, batches, type=C2C):
        """Groups of batches[g] transforms of length lengths[g], stored one group after the other
        (aclfftExtPlanManyUneven)."""
        n = len(lengths)
        if len(batches) != n:
            raise ValueError("lengths and batches differ in size")
        L = (ctypes.c_int * n)(*lengths)
        B = (ctypes.c_int * n)(*batches)
        h = ctypes.c_void_p()
        check(lib.aclfftExtPlanManyUneven(ctypes.byref(h), n, L, B, type), "aclfftExtPlanManyUneven")
        p = cls(h, type, tuple(lengths), sum(batches))
        p.groups = list(zip(lengths, batches))
        return p

    @classmethod
    def make_plan1d(cls, nx, type=C2C, batch=1, dim_type=HORIZONTAL):
        h = ctypes.c_void_p()
        check(lib.aclfftCreate(ctypes.byref(h)), "aclfftCreate")
        ws = ctypes.c_size_t(0)
        r = lib.aclfftMakePlan1d(h, nx, type, batch, dim_type, ctypes.byref(ws))
        if r != _lib.SUCCESS:
            lib.aclfftDestroy(h)
            check(r, "aclfftMakePlan1d")
        p = cls(h, type, (nx,), batch, dim_type == VERTICAL)
        p.work_size = ws.value
        return p

    # ---- management ---------------------------------------------------------------------------
    def set_stream(self, stream):
        """stream: torch.cuda.Stream, an integer hipStream_t, or None (null stream)."""
        s = getattr(stream, "cuda_stream", stream)
        check(lib.aclfftSetStream(self._h, s), "aclfftSetStream")
        self._stream = s

    def _bind_stream(self, a):
        if _is_device(a) and self._stream is None:
            import torch
            s = torch.cuda.current_stream(a.device).cuda_stream
            check(lib.aclfftSetStream(self._h, s), "aclfftSetStream")

    def destroy(self):
        if self._h is not None and self._h.value:
            check(lib.aclfftDestroy(self._h), "aclfftDestroy")
        self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---- workspace (reference: aclfftGetSize / aclfftSetAutoAllocation / aclfftSetWorkArea) ------
    def get_size(self):
        """Device workspace of one execution, in bytes (0 when the plan needs none)."""
        ws = ctypes.c_size_t(0)
        check(lib.aclfftGetSize(self._h, ctypes.byref(ws)), "aclfftGetSize")
        return ws.value

    def set_auto_allocation(self, on):
        check(lib.aclfftSetAutoAllocation(self._h, 1 if on else 0), "aclfftSetAutoAllocation")

    def set_work_area(self, area):
        """area: a device tensor of at least get_size() bytes, kept alive by the caller while the
        plan uses it."""
        if _nbytes(area) < self.get_size():
            raise ValueError("work area holds %d bytes, the plan needs %d" % (_nbytes(area), self.get_size()))
        check(lib.aclfftSetWorkArea(self._h, ctypes.c_void_p(_ptr(area))), "aclfftSetWorkArea")

    # ---- library extensions (include/ops_fft_ext.h) --------------------------------------------
    def describe(self):
        buf = ctypes.create_string_buffer(1024)
        check(lib.aclfftExtDescribe(self._h, buf, len(buf)), "aclfftExtDescribe")
        return buf.value.decode()

    def set_scale(self, scale):
        check(lib.aclfftExtSetScale(self._h, float(scale)), "aclfftExtSetScale")

    def _elem(self):
        """(real, complex) element bytes of the plan's precision."""
        return (8, 16) if self.type in (Z2Z, D2Z, Z2D) else (4, 8)

    def exec_r2c_adjoint(self, grad_out, grad_in):
        r, c = self._sizes()
        er, ec = self._elem()
        _check_size(grad_out, ec * c, "grad_out")
        _check_size(grad_in, er * r, "grad_in")
        self._bind_stream(grad_out)
        check(lib.aclfftExtExecR2CAdjoint(self._h, _ptr(grad_out), _ptr(grad_in)), "aclfftExtExecR2CAdjoint")
        return grad_in

    def exec_c2r_adjoint(self, grad_out, grad_in):
        r, c = self._sizes()
        er, ec = self._elem()
        _check_size(grad_out, er * r, "grad_out")
        _check_size(grad_in, ec * c, "grad_in")
        self._bind_stream(grad_out)
        check(lib.aclfftExtExecC2RAdjoint(self._h, _ptr(grad_out), _ptr(grad_in)), "aclfftExtExecC2RAdjoint")
        return grad_in

    def set_graph_mode(self, on):
        check(lib.aclfftExtSetGraphMode(self._h, 1 if on else 0), "aclfftExtSetGraphMode")

    def set_precision(self, mode):
        """0 FP32 Stockham; 1 MFMA BF16; 2 MFMA FP16 hi/lo; 3 MFMA BF16 hi/mid/lo; 4 automatic (chosen
        by tune())."""
        check(lib.aclfftExtSetPrecision(self._h, mode), "aclfftExtSetPrecision")

    def tune(self):
        """aclfftExtTune: measure the plan's alternatives now (on its stream: torch's current stream
        unless set_stream chose one) and keep the fastest."""
        if self._stream is None:
            import torch
            if torch.cuda.is_available():
                check(lib.aclfftSetStream(self._h, torch.cuda.current_stream().cuda_stream), "aclfftSetStream")
        check(lib.aclfftExtTune(self._h), "aclfftExtTune")

    # ---- execution ----------------------------------------------------------------------------
    def _sizes(self):
        """(real samples, complex bins) of the whole plan: the extents of the real and complex
        sides of R2C/C2R, and of both sides of C2C (complex).  Computed once per plan (the
        executions check every buffer against it)."""
        cached = self.__dict__.get("_sizes_cache")
        if cached is None:
            cached = self.__dict__["_sizes_cache"] = self._compute_sizes()
        return cached

    def _compute_sizes(self):
        if getattr(self, "groups", None):
            pts = sum(n * b for n, b in self.groups)
            return pts, pts
        full = int(np.prod(self.dims)) * self.batch
        if self.type in (C2C, Z2Z):
            return full, full
        nx = self.dims[-1]
        return full, (nx // 2 + 1) * int(np.prod(self.dims[:-1])) * self.batch

    def exec_c2c(self, idata, odata, direction=FORWARD):
        n, _ = self._sizes()
        _check_size(idata, 8 * n, "idata")
        _check_size(odata, 8 * n, "odata")
        self._bind_stream(idata)
        check(lib.aclfftExecC2C(self._h, _ptr(idata), _ptr(odata), direction), "aclfftExecC2C")
        return odata

    def exec_r2c(self, idata, odata):
        r, c = self._sizes()
        _check_size(idata, 4 * r, "idata")
        _check_size(odata, 8 * c, "odata")
        self._bind_stream(idata)
        check(lib.aclfftExecR2C(self._h, _ptr(idata), _ptr(odata)), "aclfftExecR2C")
        return odata

    def exec_c2r(self, idata, odata):
        r, c = self._sizes()
        _check_size(idata, 8 * c, "idata")
        _check_size(odata, 4 * r, "odata")
        self._bind_stream(idata)
        check(lib.aclfftExecC2R(self._h, _ptr(idata), _ptr(odata)), "aclfftExecC2R")
        return odata

    # double precision (complex128 / float64 buffers)
    def exec_z2z(self, idata, odata, direction=FORWARD):
        n, _ = self._sizes()
        _check_size(idata, 16 * n, "idata")
        _check_size(odata, 16 * n, "odata")
        self._bind_stream(idata)
        check(lib.aclfftExecZ2Z(self._h, _ptr(idata), _ptr(odata), direction), "aclfftExecZ2Z")
        return odata

    def exec_d2z(self, idata, odata):
        r, c = self._sizes()
        _check_size(idata, 8 * r, "idata")
        _check_size(odata, 16 * c, "odata")
        self._bind_stream(idata)
        check(lib.aclfftExecD2Z(self._h, _ptr(idata), _ptr(odata)), "aclfftExecD2Z")
        return odata

    def exec_z2d(self, idata, odata):
        r, c = self._sizes()
        _check_size(idata, 16 * c, "idata")
        _check_size(odata, 8 * r, "odata")
        self._bind_stream(idata)
        check(lib.aclfftExecZ2D(self._h, _ptr(idata), _ptr(odata)), "aclfftExecZ2D")
        return odata
