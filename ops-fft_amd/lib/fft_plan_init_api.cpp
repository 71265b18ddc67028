// fft_plan_init_api.cpp - aclfftMakePlan1d / aclfftMakePlan2d / aclfftMakePlan3d.
//
// Validation order and codes follow the reference lib/fft_plan_init_api.cpp:118-254 (including the
// VERTICAL encoding batch = 1, stride[0] = batch and the overflow-checked size arithmetic).  After
// validation the plan's launch list is chosen here (host only, no device call); device tables are
// created on the first execution.  *workSize receives the plan's device workspace in bytes.
#include <new>

#include "cann_ops_fft.h"
#include "executor.h"
#include "fft_handle_impl.h"
#include "fft_validate.h"

namespace {

// Byte sizes of one execution's input and output, overflow-checked (reference
// lib/fft_plan_init_api.cpp:25-105 performs the same checks and discards the values).
bool io_sizes_fit(const aclfftHandle_t* p) {
    size_t full = 1, packed = 1, t = 0;
    for (int i = 0; i < p->rank; ++i)
        if (__builtin_mul_overflow(full, (size_t)p->lengths[i], &full)) return false;
    if (__builtin_mul_overflow(full, (size_t)p->batch, &full)) return false;
    for (int i = 0; i < p->rank - 1; ++i)
        if (__builtin_mul_overflow(packed, (size_t)p->lengths[i], &packed)) return false;
    if (__builtin_mul_overflow(packed, (size_t)(p->lengths[p->rank - 1] / 2 + 1), &packed)) return false;
    if (__builtin_mul_overflow(packed, (size_t)p->batch, &packed)) return false;
    const size_t elem = (p->type == ACLFFT_Z2Z || p->type == ACLFFT_D2Z || p->type == ACLFFT_Z2D) ? 8 : 4;
    if (__builtin_mul_overflow(full, (size_t)2 * elem, &t)) return false;
    if (__builtin_mul_overflow(packed, (size_t)2 * elem, &t)) return false;
    return true;
}

aclfftResult attach_executor(aclfftHandle_t* p, const ofx::PlanDesc& d, size_t* workSize) {
    ofx::Executor* e = nullptr;
    aclfftResult r = ofx::Executor::build(d, &e);
    delete p->exec;
    p->exec = nullptr;
    if (r == ACLFFT_NOT_IMPLEMENTED) {
        // A valid plan without an execution path (e.g. a prime factor beyond the codelet set):
        // like the reference, creation succeeds and execution reports NOT_IMPLEMENTED.
        if (workSize != nullptr) *workSize = 0;
        return ACLFFT_SUCCESS;
    }
    if (r != ACLFFT_SUCCESS) return r;
    e->set_auto_allocation(p->auto_allocate != 0);
    p->exec = e;
    if (workSize != nullptr) *workSize = e->work_bytes();
    return ACLFFT_SUCCESS;
}

}  // namespace

extern "C" {

ACLFFT_API aclfftResult aclfftMakePlan1d(aclfftHandle plan, int nx, aclfftType type, int batch, int dimType,
                                         size_t* workSize) {
    if (plan == nullptr || plan->is_initialized) return ACLFFT_INVALID_VALUE;
    aclfftResult r = ofx::check_1d_args(nx, type, batch, dimType, "aclfftMakePlan1d");
    if (r != ACLFFT_SUCCESS) return r;
    plan->rank = 1;
    plan->lengths[0] = nx;
    plan->type = type;
    if (dimType == ACLFFT_VERTICAL) {
        plan->batch = 1;
        plan->stride[0] = batch;
    } else {
        plan->batch = batch;
        plan->stride[0] = 1;
    }
    if (!io_sizes_fit(plan)) return ACLFFT_INVALID_SIZE;
    ofx::PlanDesc d;
    d.rank = 1;
    d.n0 = nx;
    d.batch = batch;
    d.type = type;
    d.vertical = (dimType == ACLFFT_VERTICAL);
    r = attach_executor(plan, d, workSize);
    if (r != ACLFFT_SUCCESS) return r;
    plan->is_initialized = true;
    return ACLFFT_SUCCESS;
}

ACLFFT_API aclfftResult aclfftMakePlan2d(aclfftHandle plan, int batch, int nx, int ny, aclfftType type,
                                         size_t* workSize) {
    if (plan == nullptr || plan->is_initialized) return ACLFFT_INVALID_VALUE;
    if (nx <= 0 || ny <= 0 || batch <= 0) return ACLFFT_INVALID_SIZE;
    if (!ofx::is_known_type(type)) return ACLFFT_INVALID_TYPE;
    if (ofx::check_2d_type(type, "aclfftMakePlan2d") != ACLFFT_SUCCESS) return ACLFFT_NOT_SUPPORTED;
    aclfftResult r = ofx::check_2d_domain(nx, ny, "aclfftMakePlan2d");
    if (r != ACLFFT_SUCCESS) return r;
    plan->rank = 2;
    plan->lengths[0] = nx;
    plan->lengths[1] = ny;
    plan->batch = batch;
    plan->type = type;
    plan->stride[0] = 1;
    plan->stride[1] = 1;
    if (!io_sizes_fit(plan)) return ACLFFT_INVALID_SIZE;
    ofx::PlanDesc d;
    d.rank = 2;
    d.n0 = nx;
    d.n1 = ny;
    d.batch = batch;
    d.type = type;
    r = attach_executor(plan, d, workSize);
    if (r != ACLFFT_SUCCESS) return r;
    plan->is_initialized = true;
    return ACLFFT_SUCCESS;
}

// 3D C2C / Z2Z of one [nx][ny][nz] volume (nz fastest): rows along z, then columns along y and x.
// The reference's is a stub returning NOT_IMPLEMENTED for any arguments
// (lib/fft_plan_init_api.cpp:250-254); ACLFFT_STRICT_REFERENCE_DOMAIN=1 keeps that.
ACLFFT_API aclfftResult aclfftMakePlan3d(aclfftHandle plan, int nx, int ny, int nz, aclfftType type,
                                         size_t* workSize) {
    if (ofx::strict_reference_domain()) return ACLFFT_NOT_IMPLEMENTED;
    if (plan == nullptr || plan->is_initialized) return ACLFFT_INVALID_VALUE;
    if (nx <= 0 || ny <= 0 || nz <= 0) return ACLFFT_INVALID_SIZE;
    if (!ofx::is_known_type(type)) return ACLFFT_INVALID_TYPE;
    aclfftResult r = ofx::check_3d_type_domain(nx, ny, nz, type, "aclfftMakePlan3d");
    if (r != ACLFFT_SUCCESS) return r;
    plan->rank = 3;
    plan->lengths[0] = nx;
    plan->lengths[1] = ny;
    plan->lengths[2] = nz;
    plan->batch = 1;
    plan->type = type;
    plan->stride[0] = plan->stride[1] = plan->stride[2] = 1;
    if (!io_sizes_fit(plan)) return ACLFFT_INVALID_SIZE;
    ofx::PlanDesc d;
    d.rank = 3;
    d.n0 = nx;
    d.n1 = ny;
    d.n2 = nz;
    d.batch = 1;
    d.type = type;
    r = attach_executor(plan, d, workSize);
    if (r != ACLFFT_SUCCESS) return r;
    plan->is_initialized = true;
    return ACLFFT_SUCCESS;
}

}  // extern "C"
