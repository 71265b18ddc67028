// kernels_copy.hip - streaming device copy, the achievable-bandwidth yardstick for the transforms.
//
// One 16-byte non-temporal load and store per lane: the access pattern an HBM-bound transform
// kernel cannot beat (it reads and writes every byte once, like this copy, and does its passes in
// registers and LDS).  Measured alongside the transforms so that their fraction of the published
// HBM peak can be read against what a plain copy of the same bytes reaches.
#include <hip/hip_runtime.h>
#include "ofx_launch.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_copy(const v4f* __restrict__ src, v4f* __restrict__ dst, long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace

extern "C" OFX_API int ofx_stream_copy(const void* src, void* dst, long long bytes, hipStream_t stream) {
    if (bytes < 0 || (bytes & 15) != 0 || (((size_t)src | (size_t)dst) & 15) != 0) return hipErrorInvalidValue;
    if (bytes == 0) return hipSuccess;
    const long long n = bytes / 16;
    const long long blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<const v4f*>(src),
                       static_cast<v4f*>(dst), n);
    return hipGetLastError();
}
