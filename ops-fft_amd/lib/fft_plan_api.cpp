// fft_plan_api.cpp - aclfftCreate / aclfftPlan1d / aclfftPlan2d / aclfftPlan3d.
//
// Behaviour follows the reference lib/fft_plan_api.cpp:24-150: argument checks in the same order
// and with the same codes, then Create + MakePlan, and on a MakePlan failure the plan is destroyed
// and *plan is set to nullptr.
#include <new>

#include "cann_ops_fft.h"
#include "executor.h"
#include "fft_handle_impl.h"
#include "fft_validate.h"

aclfftHandle_t::~aclfftHandle_t() { delete exec; }

extern "C" {

ACLFFT_API aclfftResult aclfftCreate(aclfftHandle* plan) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    aclfftHandle_t* p = new (std::nothrow) aclfftHandle_t();
    if (p == nullptr) return ACLFFT_ALLOC_FAILED;
    *plan = p;
    return ACLFFT_SUCCESS;
}

ACLFFT_API aclfftResult aclfftPlan1d(aclfftHandle* plan, int nx, aclfftType type, int batch, int dimType) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    aclfftResult r = ofx::check_1d_args(nx, type, batch, dimType, "aclfftPlan1d");
    if (r != ACLFFT_SUCCESS) return r;
    r = aclfftCreate(plan);
    if (r != ACLFFT_SUCCESS) return r;
    r = aclfftMakePlan1d(*plan, nx, type, batch, dimType, nullptr);
    if (r != ACLFFT_SUCCESS) {
        aclfftDestroy(*plan);
        *plan = nullptr;
    }
    return r;
}

ACLFFT_API aclfftResult aclfftPlan2d(aclfftHandle* plan, int batch, int nx, int ny, aclfftType type) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (nx <= 0 || ny <= 0 || batch <= 0) return ACLFFT_INVALID_SIZE;
    if (!ofx::is_known_type(type)) return ACLFFT_INVALID_TYPE;
    if (ofx::check_2d_type(type, "aclfftPlan2d") != ACLFFT_SUCCESS) return ACLFFT_NOT_SUPPORTED;
    aclfftResult r = aclfftCreate(plan);
    if (r != ACLFFT_SUCCESS) return r;
    r = aclfftMakePlan2d(*plan, batch, nx, ny, type, nullptr);
    if (r != ACLFFT_SUCCESS) {
        aclfftDestroy(*plan);
        *plan = nullptr;
    }
    return r;
}

ACLFFT_API aclfftResult aclfftPlan3d(aclfftHandle* plan, int nx, int ny, int nz, aclfftType type) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (nx <= 0 || ny <= 0 || nz <= 0) return ACLFFT_INVALID_SIZE;
    if (!ofx::is_known_type(type)) return ACLFFT_INVALID_TYPE;
    aclfftResult r = aclfftCreate(plan);
    if (r != ACLFFT_SUCCESS) return r;
    r = aclfftMakePlan3d(*plan, nx, ny, nz, type, nullptr);
    if (r != ACLFFT_SUCCESS) {
        aclfftDestroy(*plan);
        *plan = nullptr;
    }
    return r;
}

}  // extern "C"
