// fft_handle_impl.h - the plan object behind aclfftHandle.
//
// The first block of fields mirrors the reference plan (lib/fft_handle_impl.h:23-59 of ops-fft):
// rank, lengths, stride, batch, type, normMode (always 0), stream and the initialised flag, with
// the same defaults and the same VERTICAL encoding (batch = 1, stride[0] = the caller's batch).
// The reference plan holds no device state; this one owns an Executor with the chosen kernels,
// the device twiddle tables, the workspace, the host-staging buffers and captured graphs.
#ifndef OFX_FFT_HANDLE_IMPL_H
#define OFX_FFT_HANDLE_IMPL_H

#include "cann_ops_fft.h"

namespace ofx {
class Executor;
}

struct aclfftHandle_t {
    int32_t rank = 0;
    int32_t lengths[3] = {0, 0, 0};
    int32_t stride[3] = {0, 0, 0};
    int32_t batch = 1;
    aclfftType type = ACLFFT_C2C;
    int32_t normMode = 0;
    hipStream_t stream = nullptr;
    bool is_initialized = false;

    // plan-owned device state, created by MakePlan (host-side choice) and filled lazily on the
    // first execution (device tables and buffers)
    ofx::Executor* exec = nullptr;
    // arithmetic selected with aclfftExtSetPrecision (0 = FP32 Stockham)
    int32_t precision = 0;
    // aclfftSetAutoAllocation: 0 = the caller provides the workspace with aclfftSetWorkArea
    int32_t auto_allocate = 1;

    aclfftHandle_t() = default;
    ~aclfftHandle_t();
    aclfftHandle_t(const aclfftHandle_t&) = delete;
    aclfftHandle_t& operator=(const aclfftHandle_t&) = delete;
};

#endif
