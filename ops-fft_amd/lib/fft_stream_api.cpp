// fft_stream_api.cpp - aclfftSetStream (reference lib/fft_stream_api.cpp:22-32): null plan ->
// INVALID_VALUE, otherwise the stream is stored (initialised or not).  The stream is a HIP stream.
#include "cann_ops_fft.h"
#include "fft_handle_impl.h"

extern "C" ACLFFT_API aclfftResult aclfftSetStream(aclfftHandle plan, hipStream_t stream) {
    if (plan == nullptr) return ACLFFT_INVALID_VALUE;
    plan->stream = stream;
    return ACLFFT_SUCCESS;
}
