"""CPU tests of the drop-in boundary: exported symbols, header ABI, and the reference's
negative-path result codes.  No GPU work is issued: every call here is rejected (or accepted)
before the library touches a device.

Expected codes come from the reference's own negative tests:
  src/fft1_d/tests/fft1_d_test_arch32.cpp:58-86,183-222 (plan gates, entry_reject_regression 6-7)
  src/fft2_d/tests/fft2_d_test_arch32.cpp:52-126 (MakePlan2d type / size gates)
  lib/fft_exec_api.cpp:32-116, lib/fft_plan_api.cpp:24-150 (exec/plan checks)
"""
import ctypes

import numpy as np
import os
import re
import subprocess

import pytest

import aclfft
from aclfft import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")

H = ctypes.c_void_p


def plan1d(nx, t, batch, dim=aclfft.HORIZONTAL):
    h = H()
    r = lib.aclfftPlan1d(ctypes.byref(h), nx, t, batch, dim)
    return r, h


def declared_functions():
    names = []
    for f in sorted(os.listdir(INC)):
        if f.endswith(".h"):
            txt = "\n".join(l for l in open(os.path.join(INC, f)).read().splitlines()
                            if not l.lstrip().startswith("#"))
            names += re.findall(r"(?:ACLFFT_API|OFX_API)\s+[\w\s\*]+?\**\s*(\w+)\s*\(", txt)
    return sorted(set(names))


def test_exports_every_declared_symbol():
    out = subprocess.check_output(["nm", "-D", "--defined-only", aclfft.LIB_PATH]).decode()
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    decl = declared_functions()
    assert set(aclfft.ACLFFT_SYMBOLS) <= set(decl)
    missing = [n for n in decl if n not in exported]
    assert not missing, missing
    # exactly the reference's 25 aclfft* names plus this library's ext/ofx additions
    acl = {n for n in exported if n.startswith("aclfft") and not n.startswith("aclfftExt")}
    assert acl == set(aclfft.ACLFFT_SYMBOLS)


def test_header_compiles_as_c_and_links(tmp_path):
    src = tmp_path / "t.c"
    src.write_text(
        '#include "cann_ops_fft.h"\n#include "ops_fft_ext.h"\n#include "ofx_launch.h"\n'
        "int main(void) { aclfftHandle p = 0; hipStream_t s = 0; aclrtStream a = s;\n"
        "  (void)a; if (aclfftPlan1d(&p, 64, ACLFFT_C2C, 1, ACLFFT_HORIZONTAL) != ACLFFT_SUCCESS) return 1;\n"
        "  if (aclfftSetStream(p, s) != ACLFFT_SUCCESS) return 2;\n"
        "  if (aclfftDestroy(p) != ACLFFT_SUCCESS) return 3;\n"
        "  return sizeof(aclfftComplex) == 8 && ACLFFT_NOT_SUPPORTED == 16 && ACLFFT_C2C == 0x29 ? 0 : 4; }\n")
    exe = tmp_path / "t"
    libdir = os.path.dirname(aclfft.LIB_PATH)
    subprocess.check_call(["gcc", "-std=c11", "-Wall", "-Werror", "-I", INC, str(src), "-o", str(exe),
                           "-L", libdir, "-lcann_ops_fft", "-Wl,-rpath," + libdir])
    assert subprocess.call([str(exe)]) == 0


def test_error_strings_match_reference():
    # lib/fft_utils_api.cpp:19-56
    expect = {0: "Success", 1: "Invalid plan handle", 2: "Memory allocation failed", 3: "Invalid data type",
              4: "Invalid parameter value", 5: "Internal error", 6: "Execution failed", 7: "Setup failed",
              8: "Invalid size parameter", 9: "Data not aligned", 10: "Incomplete parameter list",
              11: "Invalid device", 12: "Parse error", 13: "No workspace available",
              14: "Feature not implemented", 16: "Feature not supported", 15: "Unknown error", 99: "Unknown error"}
    for code, s in expect.items():
        assert aclfft.error_string(code) == s


def test_plan1d_argument_gates():
    assert lib.aclfftPlan1d(None, 64, aclfft.C2C, 1, 0) == aclfft.INVALID_VALUE
    assert plan1d(0, aclfft.C2C, 1)[0] == aclfft.INVALID_SIZE
    assert plan1d(-4, aclfft.C2C, 1)[0] == aclfft.INVALID_SIZE
    assert plan1d(64, aclfft.C2C, 0)[0] == aclfft.INVALID_SIZE
    assert plan1d(64, aclfft.C2C, 1, 2)[0] == aclfft.INVALID_VALUE
    assert plan1d(64, 0x99, 1)[0] == aclfft.INVALID_TYPE
    # issue #171 in the reference: double precision refused at plan time.  This library accepts
    # 1D double precision (a superset); the strict mode restores the reference's refusal.
    for t in (aclfft.Z2Z, aclfft.D2Z, aclfft.Z2D):
        r, h = plan1d(256, t, 1)
        assert r == aclfft.SUCCESS and h.value
        lib.aclfftDestroy(h)
    # 1D size domain (issue #190): horizontal <= 2^27; vertical pow2 in [256, 2^18]
    assert plan1d((1 << 27) + 1, aclfft.C2C, 1)[0] == aclfft.INVALID_SIZE
    for n in (128, 300, 64, (1 << 18) * 2):
        assert plan1d(n, aclfft.C2C, 1, aclfft.VERTICAL)[0] == aclfft.INVALID_SIZE
    r, h = plan1d(1 << 27, aclfft.C2C, 1)
    assert r == aclfft.SUCCESS
    lib.aclfftDestroy(h)


def test_plan1d_success_and_reuse():
    r, h = plan1d(1024, aclfft.C2C, 64)
    assert r == aclfft.SUCCESS and h.value
    ws = ctypes.c_size_t(123)
    # MakePlan on an initialised plan is INVALID_VALUE (lib/fft_error.h:124-128)
    assert lib.aclfftMakePlan1d(h, 1024, aclfft.C2C, 64, 0, ctypes.byref(ws)) == aclfft.INVALID_VALUE
    assert lib.aclfftMakePlan1d(None, 1024, aclfft.C2C, 64, 0, ctypes.byref(ws)) == aclfft.INVALID_VALUE
    assert lib.aclfftSetStream(h, None) == aclfft.SUCCESS
    assert lib.aclfftDestroy(h) == aclfft.SUCCESS
    assert lib.aclfftDestroy(None) == aclfft.INVALID_VALUE
    assert lib.aclfftSetStream(None, None) == aclfft.INVALID_VALUE


def test_makeplan_reports_workspace():
    h = H()
    assert lib.aclfftCreate(ctypes.byref(h)) == aclfft.SUCCESS
    ws = ctypes.c_size_t(7)
    assert lib.aclfftMakePlan1d(h, 1024, aclfft.C2C, 8, 0, ctypes.byref(ws)) == aclfft.SUCCESS
    assert ws.value == 0  # single-kernel plan: no workspace
    lib.aclfftDestroy(h)
    h = H()
    lib.aclfftCreate(ctypes.byref(h))
    assert lib.aclfftMakePlan1d(h, 1 << 20, aclfft.C2C, 2, 0, ctypes.byref(ws)) == aclfft.SUCCESS
    # four-step intermediate: 1024 rows of 1024 + 16 (row pitch padded off a power of two) per transform
    assert ws.value == 1024 * (1024 + 16) * 2 * 8
    lib.aclfftDestroy(h)


def test_plan2d_gates(monkeypatch):
    h = H()
    assert lib.aclfftPlan2d(None, 1, 64, 64, aclfft.C2C) == aclfft.INVALID_VALUE
    assert lib.aclfftPlan2d(ctypes.byref(h), 0, 64, 64, aclfft.C2C) == aclfft.INVALID_SIZE
    assert lib.aclfftPlan2d(ctypes.byref(h), 1, 0, 64, aclfft.C2C) == aclfft.INVALID_SIZE
    assert lib.aclfftPlan2d(ctypes.byref(h), 1, 64, -1, aclfft.C2C) == aclfft.INVALID_SIZE
    assert lib.aclfftPlan2d(ctypes.byref(h), 1, 64, 64, 0x77) == aclfft.INVALID_TYPE
    # makeplan2d_type_reject (fft2_d_test_arch32.cpp:52-86): every type but C2C in the strict
    # mode; by default 2D Z2Z is accepted as well
    for strict in (True, False):
        monkeypatch.setenv("ACLFFT_STRICT_REFERENCE_DOMAIN", "1" if strict else "0")
        refused = (aclfft.D2Z, aclfft.Z2D, aclfft.R2C, aclfft.C2R) + ((aclfft.Z2Z,) if strict else ())
        for t in refused:
            assert lib.aclfftPlan2d(ctypes.byref(h), 1, 64, 64, t) == aclfft.NOT_SUPPORTED
            p = H()
            lib.aclfftCreate(ctypes.byref(p))
            assert lib.aclfftMakePlan2d(p, 1, 64, 64, t, None) == aclfft.NOT_SUPPORTED
            lib.aclfftDestroy(p)
        ws = ctypes.c_size_t(0)
        p = H()
        lib.aclfftCreate(ctypes.byref(p))
        assert lib.aclfftGetSize2d(p, 64, 64, aclfft.Z2Z, ctypes.byref(ws)) == (
            aclfft.NOT_SUPPORTED if strict else aclfft.SUCCESS)
        lib.aclfftDestroy(p)
        accepted = (aclfft.C2C,) + (() if strict else (aclfft.Z2Z,))
        for t in accepted:
            p = H()
            lib.aclfftCreate(ctypes.byref(p))
            assert lib.aclfftMakePlan2d(p, 1, 64, 64, t, None) == aclfft.SUCCESS
            lib.aclfftDestroy(p)


@pytest.mark.parametrize("strict", [True, False])
def test_makeplan2d_size_domain(monkeypatch, strict):
    """makeplan2d_size_reject (fft2_d_test_arch32.cpp:88-126).  With
    ACLFFT_STRICT_REFERENCE_DOMAIN=1 the reference's {32,64,128}^2 gate is reproduced exactly; by
    default this library is a superset and accepts those sizes."""
    monkeypatch.setenv("ACLFFT_STRICT_REFERENCE_DOMAIN", "1" if strict else "0")
    for nx, ny in ((100, 100), (32, 100), (100, 32), (256, 256), (64, 48), (16, 32)):
        p = H()
        lib.aclfftCreate(ctypes.byref(p))
        r = lib.aclfftMakePlan2d(p, 1, nx, ny, aclfft.C2C, None)
        assert r == (aclfft.NOT_SUPPORTED if strict else aclfft.SUCCESS), (nx, ny, r)
        lib.aclfftDestroy(p)
    p = H()
    lib.aclfftCreate(ctypes.byref(p))
    assert lib.aclfftMakePlan2d(p, 1, 32, 128, aclfft.C2C, None) == aclfft.SUCCESS
    lib.aclfftDestroy(p)


def test_plan3d_gates(monkeypatch):
    """The reference's 3D entry points are stubs (NOT_IMPLEMENTED for any arguments once Plan3d's
    own checks pass, lib/fft_plan_api.cpp:127-149, lib/fft_plan_init_api.cpp:250-254,
    lib/fft_utils_api.cpp:76-79): kept in the strict mode.  By default 3D C2C / Z2Z plan, and the
    other types are NOT_SUPPORTED as in 2D."""
    h = H()
    ws = ctypes.c_size_t()
    monkeypatch.setenv("ACLFFT_STRICT_REFERENCE_DOMAIN", "1")
    assert lib.aclfftPlan3d(ctypes.byref(h), 8, 8, 8, aclfft.C2C) == aclfft.NOT_IMPLEMENTED
    assert not h.value
    assert lib.aclfftPlan3d(ctypes.byref(h), 0, 8, 8, aclfft.C2C) == aclfft.INVALID_SIZE
    assert lib.aclfftPlan3d(ctypes.byref(h), 8, 8, 8, 0x11) == aclfft.INVALID_TYPE
    assert lib.aclfftMakePlan3d(None, 8, 8, 8, aclfft.C2C, None) == aclfft.NOT_IMPLEMENTED
    assert lib.aclfftGetSize3d(None, 8, 8, 8, aclfft.C2C, ctypes.byref(ws)) == aclfft.NOT_IMPLEMENTED
    monkeypatch.setenv("ACLFFT_STRICT_REFERENCE_DOMAIN", "0")
    assert lib.aclfftPlan3d(None, 8, 8, 8, aclfft.C2C) == aclfft.INVALID_VALUE
    assert lib.aclfftPlan3d(ctypes.byref(h), 8, 0, 8, aclfft.C2C) == aclfft.INVALID_SIZE
    assert lib.aclfftPlan3d(ctypes.byref(h), 8, 8, 8, 0x11) == aclfft.INVALID_TYPE
    for t in (aclfft.R2C, aclfft.C2R, aclfft.D2Z, aclfft.Z2D):
        assert lib.aclfftPlan3d(ctypes.byref(h), 8, 8, 8, t) == aclfft.NOT_SUPPORTED and not h.value
    assert lib.aclfftPlan3d(ctypes.byref(h), 8, 1 << 28, 8, aclfft.C2C) == aclfft.NOT_SUPPORTED
    assert lib.aclfftMakePlan3d(None, 8, 8, 8, aclfft.C2C, None) == aclfft.INVALID_VALUE
    assert lib.aclfftGetSize3d(None, 8, 8, 8, aclfft.C2C, ctypes.byref(ws)) == aclfft.INVALID_PLAN
    for t in (aclfft.C2C, aclfft.Z2Z):
        assert lib.aclfftPlan3d(ctypes.byref(h), 64, 32, 16, t) == aclfft.SUCCESS and h.value
        assert lib.aclfftMakePlan3d(h, 64, 32, 16, t, None) == aclfft.INVALID_VALUE  # initialised
        lib.aclfftDestroy(h)
        p = H()
        lib.aclfftCreate(ctypes.byref(p))
        assert lib.aclfftGetSize3d(p, 64, 32, 16, t, ctypes.byref(ws)) == aclfft.SUCCESS
        assert ws.value == 0  # compiled row and column kernels in place: no workspace
        # a 2^16-point z axis runs as a four-step with a workspace the size of the data
        assert lib.aclfftGetSize3d(p, 2, 2, 1 << 16, t, ctypes.byref(ws)) == aclfft.SUCCESS
        assert ws.value >= 4 * (1 << 16) * (8 if t == aclfft.C2C else 16)
        lib.aclfftDestroy(p)


def test_placeholders():
    ws = ctypes.c_size_t()
    # the workspace API is implemented (a superset of the reference's placeholders); without a
    # plan handle it reports INVALID_PLAN
    assert lib.aclfftGetSize1d(None, 8, aclfft.C2C, 1, ctypes.byref(ws)) == aclfft.INVALID_PLAN
    assert lib.aclfftGetSize2d(None, 8, 8, aclfft.C2C, ctypes.byref(ws)) == aclfft.INVALID_PLAN
    assert lib.aclfftGetSize(None, ctypes.byref(ws)) == aclfft.INVALID_PLAN
    assert lib.aclfftSetAutoAllocation(None, 1) == aclfft.INVALID_PLAN
    assert lib.aclfftSetWorkArea(None, None) == aclfft.INVALID_PLAN
    assert lib.aclfftSetStride(None, 4) == aclfft.NOT_IMPLEMENTED
    v = ctypes.c_int()
    assert lib.aclfftGetVersion(ctypes.byref(v)) == aclfft.NOT_IMPLEMENTED
    assert lib.aclfftGetProperty(0, ctypes.byref(v)) == aclfft.NOT_IMPLEMENTED
    # double-precision execution is implemented: a null plan is INVALID_VALUE, as for C2C
    assert lib.aclfftExecZ2Z(None, None, None, -1) == aclfft.INVALID_VALUE
    assert lib.aclfftExecD2Z(None, None, None) == aclfft.INVALID_VALUE
    assert lib.aclfftExecZ2D(None, None, None) == aclfft.INVALID_VALUE


def test_exec_gates_before_any_device_call():
    buf = (ctypes.c_float * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    assert lib.aclfftExecC2C(None, p, p, -1) == aclfft.INVALID_VALUE
    h = H()
    lib.aclfftCreate(ctypes.byref(h))
    assert lib.aclfftExecC2C(h, p, p, -1) == aclfft.INVALID_PLAN  # created, not initialised
    assert lib.aclfftExecR2C(h, p, p) == aclfft.INVALID_PLAN
    lib.aclfftDestroy(h)
    r, h = plan1d(16, aclfft.C2C, 2)
    assert lib.aclfftExecC2C(h, None, p, -1) == aclfft.INVALID_VALUE
    assert lib.aclfftExecC2C(h, p, None, -1) == aclfft.INVALID_VALUE
    assert lib.aclfftExecC2C(h, p, p, 0) == aclfft.INVALID_VALUE
    assert lib.aclfftExecC2C(h, p, p, 2) == aclfft.INVALID_VALUE
    assert lib.aclfftExecR2C(h, p, p) == aclfft.INVALID_TYPE
    assert lib.aclfftExecC2R(h, p, p) == aclfft.INVALID_TYPE
    lib.aclfftDestroy(h)
    r, h = plan1d(16, aclfft.R2C, 2)
    assert lib.aclfftExecC2C(h, p, p, -1) == aclfft.INVALID_TYPE
    assert lib.aclfftExecC2R(h, p, p) == aclfft.INVALID_TYPE
    lib.aclfftDestroy(h)
    # VERTICAL R2C/C2R: valid plan, NOT_IMPLEMENTED at exec (src/common/host/fft_exec_helper.h:42-46)
    for t, fn in ((aclfft.R2C, lib.aclfftExecR2C), (aclfft.C2R, lib.aclfftExecC2R)):
        r, h = plan1d(256, t, 4, aclfft.VERTICAL)
        assert r == aclfft.SUCCESS
        assert fn(h, p, p) == aclfft.NOT_IMPLEMENTED
        lib.aclfftDestroy(h)


def test_unsupported_factor_is_not_implemented_at_exec():
    """A length with a prime factor beyond the codelet set plans successfully and reports
    NOT_IMPLEMENTED at execution, as the reference does for factors outside RADIX_MIX
    (src/fft1_d/arch35/fft1_d_exec_api.cpp:54-61)."""
    buf = (ctypes.c_float * 64)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    r, h = plan1d(10007, aclfft.C2C, 1)
    assert r == aclfft.SUCCESS
    assert lib.aclfftExecC2C(h, p, p, -1) == aclfft.NOT_IMPLEMENTED
    lib.aclfftDestroy(h)


def test_uneven_plan_argument_gates():
    import ctypes
    h = ctypes.c_void_p()
    L = (ctypes.c_int * 2)(64, 128)
    B = (ctypes.c_int * 2)(1, 2)
    lib = aclfft.lib
    assert lib.aclfftExtPlanManyUneven(None, 2, L, B, aclfft.C2C) == aclfft.INVALID_VALUE
    assert lib.aclfftExtPlanManyUneven(ctypes.byref(h), 0, L, B, aclfft.C2C) == aclfft.INVALID_SIZE
    assert lib.aclfftExtPlanManyUneven(ctypes.byref(h), 2, L, None, aclfft.C2C) == aclfft.INVALID_VALUE
    assert lib.aclfftExtPlanManyUneven(ctypes.byref(h), 2, L, B, aclfft.R2C) == aclfft.NOT_SUPPORTED
    Bz = (ctypes.c_int * 2)(1, 0)
    assert lib.aclfftExtPlanManyUneven(ctypes.byref(h), 2, L, Bz, aclfft.C2C) == aclfft.INVALID_SIZE
    assert lib.aclfftExtPlanManyUneven(ctypes.byref(h), 2, L, B, aclfft.C2C) == aclfft.SUCCESS
    # an MFMA precision on an uneven plan whose lengths have no MFMA class is refused (the plan
    # keeps its FP32 launch list)
    L3 = (ctypes.c_int * 2)(60, 243)
    h3 = ctypes.c_void_p()
    assert lib.aclfftExtPlanManyUneven(ctypes.byref(h3), 2, L3, B, aclfft.C2C) == aclfft.SUCCESS
    assert lib.aclfftExtSetPrecision(h3, aclfft.PRECISION_FP16X2) == aclfft.NOT_SUPPORTED
    assert lib.aclfftDestroy(h3) == aclfft.SUCCESS
    # nor one whose only MFMA lengths are the dense single-stage ones (no tile lists): 64 and 128
    assert lib.aclfftExtSetPrecision(h, aclfft.PRECISION_FP16X2) == aclfft.NOT_SUPPORTED
    assert lib.aclfftDestroy(h) == aclfft.SUCCESS


@pytest.mark.parametrize("shape,ok", [((4, 32, 32), True), ((4, 16, 64), True), ((4, 64, 64), True),
                                      ((4, 128, 128), False), ((4, 32, 128), False), ((4, 48, 32), False)])
def test_mfma_precision_on_2d_plans(shape, ok):
    """2D C2C plans take an MFMA precision when both axes are 16, 32 or 64 points (the launch
    list is chosen on the host, so no GPU is needed); other shapes refuse and keep FP32."""
    p = aclfft.Plan.plan2d(*shape, aclfft.C2C)
    before = p.describe()
    if ok:
        p.set_precision(aclfft.PRECISION_BF16X3)
        d = p.describe()
        assert d.startswith("mfma n=%d %dx%d" % (shape[1] * shape[2], shape[2], shape[1])), d
        assert "2d=%dx%d" % (shape[1], shape[2]) in d
    else:
        with pytest.raises(aclfft.AclfftError) as e:
            p.set_precision(aclfft.PRECISION_BF16X3)
        assert e.value.code == aclfft.NOT_SUPPORTED
        assert p.describe() == before
    p.destroy()


@pytest.mark.parametrize("n", [16, 32, 64])
def test_mfma_precision_dense_lengths(n):
    """1D C2C of 16, 32 or 64 points takes the dense single-stage MFMA kernel."""
    p = aclfft.Plan.plan1d(n, aclfft.C2C, 8)
    p.set_precision(aclfft.PRECISION_FP16X2)
    assert p.describe().startswith("mfma n=%d 1x%d" % (n, n)), p.describe()
    p.destroy()


def test_python_plan_refuses_short_buffers():
    """The C entry points trust the caller's buffer extents; the Python layer checks them before
    any device work (no GPU needed: the check fires before the library call)."""
    c64 = np.complex64
    p = aclfft.Plan.plan1d(1024, aclfft.C2C, 4)
    with pytest.raises(ValueError):
        p.exec_c2c(np.zeros(4 * 1024 - 1, c64), np.zeros(4 * 1024, c64))
    with pytest.raises(ValueError):
        p.exec_c2c(np.zeros(4 * 1024, c64), np.zeros(4 * 1024 - 1, c64))
    p.destroy()
    p = aclfft.Plan.plan1d(4096, aclfft.R2C, 3)
    with pytest.raises(ValueError):
        p.exec_r2c(np.zeros(3 * 4096 - 1, np.float32), np.zeros(3 * 2049, c64))
    with pytest.raises(ValueError):
        p.exec_r2c(np.zeros(3 * 4096, np.float32), np.zeros(3 * 2049 - 1, c64))
    p.destroy()
    p = aclfft.Plan.plan1d(3000, aclfft.C2R, 2)
    with pytest.raises(ValueError):
        p.exec_c2r(np.zeros(2 * 1501 - 1, c64), np.zeros(2 * 3000, np.float32))
    with pytest.raises(ValueError):
        p.exec_c2r_adjoint(np.zeros(2 * 3000 - 1, np.float32), np.zeros(2 * 1501, c64))
    p.destroy()
    p = aclfft.Plan.plan2d(2, 64, 32, aclfft.C2C)
    with pytest.raises(ValueError):
        p.exec_c2c(np.zeros(2 * 64 * 32 - 1, c64), np.zeros(2 * 64 * 32, c64))
    p.destroy()
    p = aclfft.Plan.plan_many_uneven([60, 1000], [3, 2])
    with pytest.raises(ValueError):
        p.exec_c2c(np.zeros(60 * 3 + 1000 * 2 - 1, c64), np.zeros(2180, c64))
    p.destroy()


def test_workspace_api_host_side():
    """aclfftGetSize* agree with MakePlan's workSize; argument checks follow MakePlan's codes;
    a work area must be device memory (checked before it is used)."""
    H = ctypes.c_void_p
    h = H()
    assert lib.aclfftCreate(ctypes.byref(h)) == aclfft.SUCCESS
    ws = ctypes.c_size_t(123)
    for n, t, b in ((1024, aclfft.C2C, 8), (1 << 20, aclfft.C2C, 2), (1 << 22, aclfft.C2C, 1),
                    (3000, aclfft.R2C, 4), (1 << 21, aclfft.R2C, 1), (1 << 21, aclfft.C2R, 3)):
        assert lib.aclfftGetSize1d(h, n, t, b, ctypes.byref(ws)) == aclfft.SUCCESS
        est = ws.value
        h2 = H()
        assert lib.aclfftCreate(ctypes.byref(h2)) == aclfft.SUCCESS
        assert lib.aclfftMakePlan1d(h2, n, t, b, 0, ctypes.byref(ws)) == aclfft.SUCCESS
        assert ws.value == est, (n, t, b, ws.value, est)
        got = ctypes.c_size_t(0)
        assert lib.aclfftGetSize(h2, ctypes.byref(got)) == aclfft.SUCCESS and got.value == est
        lib.aclfftDestroy(h2)
    assert lib.aclfftGetSize1d(h, 1024, aclfft.C2C, 8, ctypes.byref(ws)) == aclfft.SUCCESS and ws.value == 0
    # real plans also hold the input copy an in-place execution on device memory reads from
    assert lib.aclfftGetSize1d(h, 4096, aclfft.R2C, 2, ctypes.byref(ws)) == aclfft.SUCCESS
    assert ws.value >= 4096 * 2 * 4
    # MakePlan's argument codes
    assert lib.aclfftGetSize1d(h, 0, aclfft.C2C, 1, ctypes.byref(ws)) == aclfft.INVALID_SIZE
    assert lib.aclfftGetSize1d(h, 64, 0x7, 1, ctypes.byref(ws)) == aclfft.INVALID_TYPE
    assert lib.aclfftGetSize1d(h, 64, aclfft.C2C, 1, None) == aclfft.INVALID_VALUE
    assert lib.aclfftGetSize2d(h, 2048, 2048, aclfft.C2C, ctypes.byref(ws)) == aclfft.SUCCESS
    assert lib.aclfftGetSize2d(h, 64, 64, aclfft.R2C, ctypes.byref(ws)) == aclfft.NOT_SUPPORTED
    assert lib.aclfftGetSize2d(h, 0, 64, aclfft.C2C, ctypes.byref(ws)) == aclfft.INVALID_SIZE
    # GetSize / SetWorkArea need a made plan; SetAutoAllocation does not
    assert lib.aclfftGetSize(h, ctypes.byref(ws)) == aclfft.INVALID_PLAN
    assert lib.aclfftSetWorkArea(h, ctypes.c_void_p(4096)) == aclfft.INVALID_PLAN
    assert lib.aclfftSetAutoAllocation(h, 0) == aclfft.SUCCESS
    assert lib.aclfftMakePlan1d(h, 1 << 20, aclfft.C2C, 1, 0, ctypes.byref(ws)) == aclfft.SUCCESS
    assert ws.value > 0
    assert lib.aclfftSetWorkArea(h, None) == aclfft.INVALID_VALUE
    host = (ctypes.c_char * ws.value)()
    assert lib.aclfftSetWorkArea(h, ctypes.cast(host, ctypes.c_void_p)) == aclfft.INVALID_VALUE
    lib.aclfftDestroy(h)


def test_double_precision_gates(monkeypatch):
    """1D Z2Z / D2Z / Z2D plans: accepted by default, refused (reference issue #171) in the strict
    mode; exec type and direction checks as for the single-precision entry points."""
    H = ctypes.c_void_p
    monkeypatch.setenv("ACLFFT_STRICT_REFERENCE_DOMAIN", "1")
    for t in (aclfft.Z2Z, aclfft.D2Z, aclfft.Z2D):
        h = H()
        assert lib.aclfftPlan1d(ctypes.byref(h), 256, t, 1, 0) == aclfft.NOT_SUPPORTED and not h.value
    monkeypatch.delenv("ACLFFT_STRICT_REFERENCE_DOMAIN")
    buf = (ctypes.c_double * 1024)()
    p = ctypes.cast(buf, ctypes.c_void_p)
    h = H()
    assert lib.aclfftPlan1d(ctypes.byref(h), 256, aclfft.Z2Z, 2, 0) == aclfft.SUCCESS
    assert lib.aclfftExecZ2Z(h, p, p, 0) == aclfft.INVALID_VALUE        # direction
    assert lib.aclfftExecC2C(h, p, p, -1) == aclfft.INVALID_TYPE        # plan/exec type mismatch
    assert lib.aclfftExecD2Z(h, p, p) == aclfft.INVALID_TYPE
    assert lib.aclfftExecZ2Z(h, None, p, -1) == aclfft.INVALID_VALUE
    lib.aclfftDestroy(h)
    h = H()
    assert lib.aclfftPlan1d(ctypes.byref(h), 512, aclfft.D2Z, 1, aclfft.VERTICAL) == aclfft.SUCCESS
    assert lib.aclfftExecD2Z(h, p, p) == aclfft.NOT_IMPLEMENTED         # VERTICAL real layout
    lib.aclfftDestroy(h)
    h = H()
    assert lib.aclfftPlan1d(ctypes.byref(h), 4096, aclfft.Z2D, 3, 0) == aclfft.SUCCESS
    assert lib.aclfftExecZ2D(h, p, p) != aclfft.INVALID_TYPE
    lib.aclfftDestroy(h)


def test_tuning_signatures_tell_every_kernel_apart():
    """The tuning cache records a step's choice by the kernel's configuration (kind, length, lanes,
    transforms per block, staging, I/O width, radices, streaming, LDS footprint), not by registry
    index: no two compiled kernels of either registry may share one, or a cached choice could
    come back as another kernel."""
    from aclfft._lib import KernelInfo

    def sig(k):
        return (k.kind, k.n, k.threads_per_transform, k.transforms_per_block, k.staged, k.iov,
                tuple(k.radices[:k.npass]), k.stream, k.lds_bytes)

    for count, get in ((lib.ofx_kernel_count, lib.ofx_kernel_info_get),
                       (lib.ofx_col_kernel_count, lib.ofx_col_kernel_info_get)):
        seen = {}
        for i in range(count()):
            k = KernelInfo()
            assert get(i, ctypes.byref(k)) == 0
            s = sig(k)
            assert s not in seen, (i, seen.get(s), s)
            seen[s] = i


@pytest.mark.parametrize("n,batch,streamed", [(8192, 1024, True), (4096, 2048, True), (4096, 8192, False),
                                              (3000, 8192, False)])
def test_r2c_store_rule_by_output_size(n, batch, streamed):
    """R2C plans of one launch stream their stores from 16 MiB up to 64 MiB of output; larger ones
    keep cached stores (the launch list is chosen on the host, so no GPU is needed)."""
    p = aclfft.Plan.plan1d(n, aclfft.R2C, batch)
    assert ("ntstore" in p.describe()) == streamed, p.describe()
    p.destroy()
