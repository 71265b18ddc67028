// fft_ext_api.cpp - library additions declared in include/ops_fft_ext.h.
#include <cstring>

#include "executor.h"
#include "fft_handle_impl.h"
#include "fft_validate.h"
#include "ops_fft_ext.h"

extern "C" {

ACLFFT_API aclfftResult aclfftExtDescribe(aclfftHandle plan, char* buf, size_t len) {
    if (plan == nullptr || buf == nullptr || len == 0) return ACLFFT_INVALID_VALUE;
    if (!plan->is_initialized) return ACLFFT_INVALID_PLAN;
    const std::string d = plan->exec ? plan->exec->describe() : std::string("not implemented");
    std::strncpy(buf, d.c_str(), len - 1);
    buf[len - 1] = '\0';
    return ACLFFT_SUCCESS;
}

ACLFFT_API aclfftResult aclfftExtSetPrecision(aclfftHandle plan, int precision) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (!plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (precision < ACLFFT_EXT_PRECISION_FP32 || precision > ACLFFT_EXT_PRECISION_AUTO) return ACLFFT_INVALID_VALUE;
    if (plan->exec == nullptr) return ACLFFT_NOT_SUPPORTED;
    ofx::PlanDesc d = plan->exec->desc();
    d.auto_precision = precision == ACLFFT_EXT_PRECISION_AUTO;
    d.precision = d.auto_precision ? ACLFFT_EXT_PRECISION_FP32 : precision;
    if (d.auto_precision) {
        // automatic precision needs an MFMA alternative to choose from
        ofx::PlanDesc probe = d;
        probe.precision = ACLFFT_EXT_PRECISION_FP16X2;
        ofx::Executor* t = nullptr;
        const aclfftResult rp = ofx::Executor::build(probe, &t);
        delete t;
        if (rp != ACLFFT_SUCCESS) return rp == ACLFFT_NOT_IMPLEMENTED ? ACLFFT_NOT_SUPPORTED : rp;
    }
    ofx::Executor* e = nullptr;
    aclfftResult r = ofx::Executor::build(d, &e);
    if (r != ACLFFT_SUCCESS) return r == ACLFFT_NOT_IMPLEMENTED ? ACLFFT_NOT_SUPPORTED : r;
    e->set_graphs(plan->exec->graphs());
    e->inherit_workspace(*plan->exec);
    delete plan->exec;
    plan->exec = e;
    plan->precision = precision;
    return ACLFFT_SUCCESS;
}

ACLFFT_API aclfftResult aclfftExtPlanManyUneven(aclfftHandle* plan, int count, const int* lengths, const int* batches,
                                               aclfftType type) {
    if (plan == nullptr || lengths == nullptr || batches == nullptr) return ACLFFT_INVALID_VALUE;
    if (count <= 0) return ACLFFT_INVALID_SIZE;
    if (!ofx::is_known_type(type)) return ACLFFT_INVALID_TYPE;
    if (type != ACLFFT_C2C) return ACLFFT_NOT_SUPPORTED;
    size_t total = 0;
    ofx::PlanDesc d;
    d.rank = 1;
    d.type = type;
    d.uneven = true;
    for (int g = 0; g < count; ++g) {
        if (lengths[g] <= 0 || lengths[g] > ofx::kMaxHorizontalN || batches[g] <= 0) return ACLFFT_INVALID_SIZE;
        size_t e = 0;
        if (__builtin_mul_overflow((size_t)lengths[g], (size_t)batches[g], &e) ||
            __builtin_add_overflow(total, e, &total))
            return ACLFFT_INVALID_SIZE;
        d.ulen.push_back(lengths[g]);
        d.ubatch.push_back(batches[g]);
    }
    d.batch = count;
    aclfftResult r = aclfftCreate(plan);
    if (r != ACLFFT_SUCCESS) return r;
    ofx::Executor* e = nullptr;
    r = ofx::Executor::build(d, &e);
    if (r != ACLFFT_SUCCESS) {
        aclfftDestroy(*plan);
        *plan = nullptr;
        return r;
    }
    aclfftHandle_t* p = *plan;
    p->rank = 1;
    p->lengths[0] = 0;
    p->batch = count;
    p->stride[0] = 1;
    p->type = type;
    p->exec = e;
    p->is_initialized = true;
    return ACLFFT_SUCCESS;
}

ACLFFT_API aclfftResult aclfftExtSetScale(aclfftHandle plan, float scale) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (!plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (!(scale == scale) || scale == 0.f) return ACLFFT_INVALID_VALUE;
    if (plan->exec == nullptr) return ACLFFT_NOT_SUPPORTED;
    ofx::PlanDesc d = plan->exec->desc();
    d.scale = scale;
    ofx::Executor* e = nullptr;
    aclfftResult r = ofx::Executor::build(d, &e);
    if (r != ACLFFT_SUCCESS) return r;
    e->set_graphs(plan->exec->graphs());
    e->inherit_workspace(*plan->exec);
    delete plan->exec;
    plan->exec = e;
    return ACLFFT_SUCCESS;
}

static aclfftResult exec_adjoint(aclfftHandle plan, const void* g, void* out, aclfftType expect) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (!plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (g == nullptr || out == nullptr) return ACLFFT_INVALID_VALUE;
    // the real-to-complex adjoint serves R2C and D2Z plans, the complex-to-real one C2R and Z2D
    const aclfftType twin = expect == ACLFFT_R2C ? ACLFFT_D2Z : ACLFFT_Z2D;
    if (plan->type != expect && plan->type != twin) return ACLFFT_INVALID_TYPE;
    if (plan->rank != 1 || plan->stride[0] > 1 || plan->exec == nullptr) return ACLFFT_NOT_IMPLEMENTED;
    ofx::Executor* a = plan->exec->adjoint();
    if (a == nullptr) return ACLFFT_NOT_IMPLEMENTED;
    return a->exec(g, out, expect == ACLFFT_R2C ? ACLFFT_BACKWARD : ACLFFT_FORWARD, plan->stream);
}

ACLFFT_API aclfftResult aclfftExtExecR2CAdjoint(aclfftHandle plan, aclfftComplex* grad_out, aclfftReal* grad_in) {
    return exec_adjoint(plan, grad_out, grad_in, ACLFFT_R2C);
}

ACLFFT_API aclfftResult aclfftExtExecC2RAdjoint(aclfftHandle plan, aclfftReal* grad_out, aclfftComplex* grad_in) {
    return exec_adjoint(plan, grad_out, grad_in, ACLFFT_C2R);
}

ACLFFT_API aclfftResult aclfftExtExecC2CNoAlloc(aclfftHandle plan, aclfftComplex* idata, aclfftComplex* odata,
                                                int direction) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (!plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (idata == nullptr || odata == nullptr) return ACLFFT_INVALID_VALUE;
    if (plan->type != ACLFFT_C2C) return ACLFFT_INVALID_TYPE;
    if (direction != ACLFFT_FORWARD && direction != ACLFFT_BACKWARD) return ACLFFT_INVALID_VALUE;
    if (plan->exec == nullptr) return ACLFFT_NOT_IMPLEMENTED;
    return plan->exec->exec_noalloc(idata, odata, direction, plan->stream);
}

ACLFFT_API aclfftResult aclfftExtTune(aclfftHandle plan) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (!plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (plan->exec == nullptr) return ACLFFT_NOT_IMPLEMENTED;
    return plan->exec->tune(plan->stream);
}

ACLFFT_API aclfftResult aclfftExtSetGraphMode(aclfftHandle plan, int on) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    if (!plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (plan->exec == nullptr) return ACLFFT_NOT_SUPPORTED;
    plan->exec->set_graphs(on != 0);
    return ACLFFT_SUCCESS;
}

}  // extern "C"
