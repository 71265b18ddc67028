// fft_exec_api.cpp - aclfftExecC2C / aclfftExecR2C / aclfftExecC2R and their double-precision
// counterparts aclfftExecZ2Z / aclfftExecD2Z / aclfftExecZ2D.
//
// Checks and codes follow the reference lib/fft_exec_api.cpp:32-116 and the operator-level entry
// checks (src/common/host/fft_exec_helper.h:37-61; src/fft1_d/arch35/fft1_d_exec_api.cpp:19-27):
// null plan -> INVALID_VALUE, uninitialised -> INVALID_PLAN, null data -> INVALID_VALUE, plan/exec
// type mismatch -> INVALID_TYPE, C2C direction not +-1 -> INVALID_VALUE.  R2C/C2R on a VERTICAL
// plan -> NOT_IMPLEMENTED (reference fft_exec_helper.h:42-46).
#include "cann_ops_fft.h"
#include "executor.h"
#include "fft_handle_impl.h"
#include "fft_validate.h"

namespace {

aclfftResult common_checks(aclfftHandle plan, const void* idata, const void* odata, aclfftType expect) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (!plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (idata == nullptr || odata == nullptr) return ACLFFT_INVALID_VALUE;
    if (plan->type != expect) return ACLFFT_INVALID_TYPE;
    return ACLFFT_SUCCESS;
}

aclfftResult run(aclfftHandle plan, const void* in, void* out, int direction) {
    if (plan->exec == nullptr) return ACLFFT_NOT_IMPLEMENTED;
    return plan->exec->exec(in, out, direction, plan->stream);
}

}  // namespace

extern "C" {

ACLFFT_API aclfftResult aclfftExecC2C(aclfftHandle plan, aclfftComplex* idata, aclfftComplex* odata, int direction) {
    aclfftResult r = common_checks(plan, idata, odata, ACLFFT_C2C);
    if (r != ACLFFT_SUCCESS) return r;
    if (direction != ACLFFT_FORWARD && direction != ACLFFT_BACKWARD) return ACLFFT_INVALID_VALUE;
    if (plan->rank < 1 || plan->rank > 3) return ACLFFT_NOT_IMPLEMENTED;
    return run(plan, idata, odata, direction);
}

ACLFFT_API aclfftResult aclfftExecR2C(aclfftHandle plan, aclfftReal* idata, aclfftComplex* odata) {
    aclfftResult r = common_checks(plan, idata, odata, ACLFFT_R2C);
    if (r != ACLFFT_SUCCESS) return r;
    if (plan->rank != 1) return ACLFFT_NOT_IMPLEMENTED;
    if (plan->stride[0] > 1) {
        ofx::log_reject("R2C", "VERTICAL layout is not supported");
        return ACLFFT_NOT_IMPLEMENTED;
    }
    return run(plan, idata, odata, ACLFFT_FORWARD);
}

ACLFFT_API aclfftResult aclfftExecC2R(aclfftHandle plan, aclfftComplex* idata, aclfftReal* odata) {
    aclfftResult r = common_checks(plan, idata, odata, ACLFFT_C2R);
    if (r != ACLFFT_SUCCESS) return r;
    if (plan->rank != 1) return ACLFFT_NOT_IMPLEMENTED;
    if (plan->stride[0] > 1) {
        ofx::log_reject("C2R", "VERTICAL layout is not supported");
        return ACLFFT_NOT_IMPLEMENTED;
    }
    return run(plan, idata, odata, ACLFFT_BACKWARD);
}

// Double precision.  The reference declares these (include/cann_ops_fft.h) and ships them as
// NOT_IMPLEMENTED stubs (lib/fft_exec_api.cpp:87-116), refusing D2Z/Z2D/Z2Z plans at creation; here
// double-precision 1D plans (HORIZONTAL, and VERTICAL for Z2Z) and 2D Z2Z plans are accepted
// (ACLFFT_STRICT_REFERENCE_DOMAIN=1 restores the refusal) and these entry points run them with the
// same checks as their single-precision counterparts.
ACLFFT_API aclfftResult aclfftExecZ2Z(aclfftHandle plan, aclfftDoubleComplex* idata, aclfftDoubleComplex* odata,
                                      int direction) {
    aclfftResult r = common_checks(plan, idata, odata, ACLFFT_Z2Z);
    if (r != ACLFFT_SUCCESS) return r;
    if (direction != ACLFFT_FORWARD && direction != ACLFFT_BACKWARD) return ACLFFT_INVALID_VALUE;
    if (plan->rank < 1 || plan->rank > 3) return ACLFFT_NOT_IMPLEMENTED;
    return run(plan, idata, odata, direction);
}

ACLFFT_API aclfftResult aclfftExecD2Z(aclfftHandle plan, aclfftDoubleReal* idata, aclfftDoubleComplex* odata) {
    aclfftResult r = common_checks(plan, idata, odata, ACLFFT_D2Z);
    if (r != ACLFFT_SUCCESS) return r;
    if (plan->rank != 1) return ACLFFT_NOT_IMPLEMENTED;
    if (plan->stride[0] > 1) {
        ofx::log_reject("D2Z", "VERTICAL layout is not supported");
        return ACLFFT_NOT_IMPLEMENTED;
    }
    return run(plan, idata, odata, ACLFFT_FORWARD);
}

ACLFFT_API aclfftResult aclfftExecZ2D(aclfftHandle plan, aclfftDoubleComplex* idata, aclfftDoubleReal* odata) {
    aclfftResult r = common_checks(plan, idata, odata, ACLFFT_Z2D);
    if (r != ACLFFT_SUCCESS) return r;
    if (plan->rank != 1) return ACLFFT_NOT_IMPLEMENTED;
    if (plan->stride[0] > 1) {
        ofx::log_reject("Z2D", "VERTICAL layout is not supported");
        return ACLFFT_NOT_IMPLEMENTED;
    }
    return run(plan, idata, odata, ACLFFT_BACKWARD);
}

}  // extern "C"
