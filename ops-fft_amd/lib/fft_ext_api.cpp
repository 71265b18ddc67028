// fft_ext_api.cpp - library additions declared in include/ops_fft_ext.h.
#include <cstring>

#include "executor.h"
#include "fft_handle_impl.h"
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

}  // extern "C"
