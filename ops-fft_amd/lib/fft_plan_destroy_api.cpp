// fft_plan_destroy_api.cpp - aclfftDestroy (reference lib/fft_plan_destroy_api.cpp:24-40): null ->
// INVALID_VALUE; otherwise the plan and every device buffer it owns are released.  Destroying a
// plan twice is undefined, as in the reference.
#include "cann_ops_fft.h"
#include "fft_handle_impl.h"

extern "C" ACLFFT_API aclfftResult aclfftDestroy(aclfftHandle plan) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    delete plan;
    return ACLFFT_SUCCESS;
}
