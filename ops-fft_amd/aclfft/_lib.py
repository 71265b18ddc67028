"""ctypes binding of libcann_ops_fft.so (the aclfft C API and the ofx kernel C-ABI).

The library is loaded from the package's parent directory (built in-tree by `make -C ops-fft_amd`).
There is no fallback: if the library is missing, importing this module raises.
"""
import ctypes
import os

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# ACLFFT_LIBRARY selects another build of the library (the sanitizer build of the host layer)
LIB_PATH = os.environ.get("ACLFFT_LIBRARY") or os.path.join(os.path.dirname(_PKG_DIR), "libcann_ops_fft.so")

# aclfftResult values (include/cann_ops_fft.h)
SUCCESS = 0
INVALID_PLAN = 1
ALLOC_FAILED = 2
INVALID_TYPE = 3
INVALID_VALUE = 4
INTERNAL_ERROR = 5
EXEC_FAILED = 6
SETUP_FAILED = 7
INVALID_SIZE = 8
UNALIGNED_DATA = 9
INCOMPLETE_PARAMETER_LIST = 10
INVALID_DEVICE = 11
PARSE_ERROR = 12
NO_WORKSPACE = 13
NOT_IMPLEMENTED = 14
NOT_SUPPORTED = 16

# aclfftType
R2C = 0x2A
C2R = 0x2C
C2C = 0x29
D2Z = 0x6A
Z2D = 0x6C
Z2Z = 0x69

# aclfftExtSetPrecision modes (include/ops_fft_ext.h)
PRECISION_FP32 = 0
PRECISION_BF16 = 1
PRECISION_FP16X2 = 2
PRECISION_BF16X3 = 3
PRECISION_AUTO = 4

FORWARD = -1
BACKWARD = 1
HORIZONTAL = 0
VERTICAL = 1

RESULT_NAMES = {
    0: "ACLFFT_SUCCESS", 1: "ACLFFT_INVALID_PLAN", 2: "ACLFFT_ALLOC_FAILED", 3: "ACLFFT_INVALID_TYPE",
    4: "ACLFFT_INVALID_VALUE", 5: "ACLFFT_INTERNAL_ERROR", 6: "ACLFFT_EXEC_FAILED", 7: "ACLFFT_SETUP_FAILED",
    8: "ACLFFT_INVALID_SIZE", 9: "ACLFFT_UNALIGNED_DATA", 10: "ACLFFT_INCOMPLETE_PARAMETER_LIST",
    11: "ACLFFT_INVALID_DEVICE", 12: "ACLFFT_PARSE_ERROR", 13: "ACLFFT_NO_WORKSPACE",
    14: "ACLFFT_NOT_IMPLEMENTED", 16: "ACLFFT_NOT_SUPPORTED",
}

# names the library must export (the reference's 25 aclfft entry points)
ACLFFT_SYMBOLS = (
    "aclfftPlan1d", "aclfftPlan2d", "aclfftPlan3d", "aclfftCreate", "aclfftMakePlan1d", "aclfftMakePlan2d",
    "aclfftMakePlan3d", "aclfftGetSize1d", "aclfftGetSize2d", "aclfftGetSize3d", "aclfftGetSize",
    "aclfftSetAutoAllocation", "aclfftSetWorkArea", "aclfftExecC2C", "aclfftExecR2C", "aclfftExecC2R",
    "aclfftExecZ2Z", "aclfftExecD2Z", "aclfftExecZ2D", "aclfftSetStream", "aclfftSetStride", "aclfftDestroy",
    "aclfftGetVersion", "aclfftGetProperty", "aclfftGetErrorString",
)


class AclfftError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__("%s%s (%d: %s)" % (what + ": " if what else "", RESULT_NAMES.get(code, "?"), code,
                                            error_string(code)))


class KernelInfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("threads_per_transform", ctypes.c_int),
                ("transforms_per_block", ctypes.c_int), ("block_threads", ctypes.c_int), ("npass", ctypes.c_int),
                ("radices", ctypes.c_int * 12), ("staged", ctypes.c_int), ("lds_bytes", ctypes.c_int),
                ("tw_len", ctypes.c_int), ("iov", ctypes.c_int), ("stream", ctypes.c_int)]


# One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7, and its libraries
# link it as "libamdhip64.so", so if this library were loaded first (resolving the system
# /opt/rocm copy) a later `import torch` would load a second runtime whose device pointers and
# streams this library could not see.  Loading torch first makes libcann_ops_fft.so bind to the
# runtime torch already loaded (same soname).
try:
    import torch as _torch  # noqa: F401
except ImportError:  # C-only users: the system ROCm runtime
    _torch = None

if not os.path.exists(LIB_PATH):
    raise ImportError("libcann_ops_fft.so not built at %s (run `make -C ops-fft_amd` or "
                      "__graft_entry__.build())" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

_h = ctypes.c_void_p
_hp = ctypes.POINTER(ctypes.c_void_p)
_i = ctypes.c_int
_sz = ctypes.POINTER(ctypes.c_size_t)
_vp = ctypes.c_void_p

_SIGS = {
    "aclfftPlan1d": [_hp, _i, _i, _i, _i],
    "aclfftPlan2d": [_hp, _i, _i, _i, _i],
    "aclfftPlan3d": [_hp, _i, _i, _i, _i],
    "aclfftCreate": [_hp],
    "aclfftMakePlan1d": [_h, _i, _i, _i, _i, _sz],
    "aclfftMakePlan2d": [_h, _i, _i, _i, _i, _sz],
    "aclfftMakePlan3d": [_h, _i, _i, _i, _i, _sz],
    "aclfftGetSize1d": [_h, _i, _i, _i, _sz],
    "aclfftGetSize2d": [_h, _i, _i, _i, _sz],
    "aclfftGetSize3d": [_h, _i, _i, _i, _i, _sz],
    "aclfftGetSize": [_h, _sz],
    "aclfftSetAutoAllocation": [_h, _i],
    "aclfftSetWorkArea": [_h, _vp],
    "aclfftExecC2C": [_h, _vp, _vp, _i],
    "aclfftExecR2C": [_h, _vp, _vp],
    "aclfftExecC2R": [_h, _vp, _vp],
    "aclfftExecZ2Z": [_h, _vp, _vp, _i],
    "aclfftExecD2Z": [_h, _vp, _vp],
    "aclfftExecZ2D": [_h, _vp, _vp],
    "aclfftSetStream": [_h, _vp],
    "aclfftSetStride": [_h, ctypes.c_int32],
    "aclfftDestroy": [_h],
    "aclfftGetVersion": [ctypes.POINTER(_i)],
    "aclfftGetProperty": [_i, ctypes.POINTER(_i)],
}
for _name, _args in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _i
lib.aclfftExtDescribe.argtypes = [_h, ctypes.c_char_p, ctypes.c_size_t]
lib.aclfftExtDescribe.restype = _i
lib.aclfftExtPlanManyUneven.argtypes = [_hp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), _i]
lib.aclfftExtPlanManyUneven.restype = _i
lib.aclfftExtSetGraphMode.argtypes = [_h, _i]
lib.aclfftExtSetGraphMode.restype = _i
lib.aclfftExtSetScale.argtypes = [_h, ctypes.c_float]
lib.aclfftExtSetScale.restype = _i
lib.aclfftExtExecR2CAdjoint.argtypes = [_h, _vp, _vp]
lib.aclfftExtExecR2CAdjoint.restype = _i
lib.aclfftExtExecC2RAdjoint.argtypes = [_h, _vp, _vp]
lib.aclfftExtExecC2RAdjoint.restype = _i
lib.aclfftExtSetPrecision.argtypes = [_h, _i]
lib.aclfftExtSetPrecision.restype = _i
lib.aclfftGetErrorString.argtypes = [_i]
lib.aclfftGetErrorString.restype = ctypes.c_char_p
lib.ofx_kernel_count.argtypes = []
lib.ofx_kernel_count.restype = _i
lib.ofx_kernel_info_get.argtypes = [_i, ctypes.POINTER(KernelInfo)]
lib.ofx_kernel_info_get.restype = _i
lib.ofx_stream_copy.argtypes = [_vp, _vp, ctypes.c_longlong, _vp]
lib.ofx_stream_copy.restype = _i


def error_string(code):
    return lib.aclfftGetErrorString(code).decode()


def check(code, what=""):
    if code != SUCCESS:
        raise AclfftError(code, what)


def kernel_registry():
    out = []
    for i in range(lib.ofx_kernel_count()):
        k = KernelInfo()
        lib.ofx_kernel_info_get(i, ctypes.byref(k))
        out.append(dict(id=i, kind=k.kind, n=k.n, T=k.threads_per_transform, TPB=k.transforms_per_block,
                        block=k.block_threads, radices=list(k.radices[:k.npass]), staged=bool(k.staged),
                        lds_bytes=k.lds_bytes, tw_len=k.tw_len, iov=k.iov, stream=int(k.stream)))
    return out


def stream_copy(src, dst, nbytes, stream=None):
    """Streaming device copy of `nbytes` (a multiple of 16) between 16-byte aligned device
    addresses on `stream` (a hipStream_t handle or None for the null stream)."""
    code = lib.ofx_stream_copy(ctypes.c_void_p(src), ctypes.c_void_p(dst), nbytes, stream)
    if code != 0:
        raise RuntimeError("ofx_stream_copy failed: hip error %d" % code)
