// twiddle64.h - four-step twiddles W_N^e for N up to 2^27, formed in FP64 and rounded once.
//
// Table (generated on the GPU with FP64 sincospi, ofx_gen_ft_table): t[e] = W_N^e for e < B and
// t[B + h] = W_N^(h*B) for h < ceil(N/B), B = 2048.  Then W_N^e = t[e mod B] * t[B + e div B] is
// one FP64 complex product of two correctly rounded values (error ~1e-16), rounded to FP32.
#pragma once
#include <hip/hip_runtime.h>

namespace ofx {

constexpr int kFtLoBits = 11;  // B = 2048
constexpr long long kFtLo = 1LL << kFtLoBits;

// e must lie in [0, len) (the four-step products column * row index do).
__device__ __forceinline__ float2 ft_twiddle(const double2* __restrict__ t, long long e, long long len) {
    (void)len;
    const long long lo = e & (kFtLo - 1);
    const long long hi = e >> kFtLoBits;
    const double2 a = t[lo];
    const double2 b = t[kFtLo + hi];
    return make_float2((float)(a.x * b.x - a.y * b.y), (float)(a.x * b.y + a.y * b.x));
}

// the same twiddle kept in FP64 (complex128 kernels)
__device__ __forceinline__ double2 ft_twiddle_z(const double2* __restrict__ t, long long e) {
    const double2 a = t[e & (kFtLo - 1)];
    const double2 b = t[kFtLo + (e >> kFtLoBits)];
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 ft_twiddle_v(float2*, const double2* __restrict__ t, long long e, long long len) {
    return ft_twiddle(t, e, len);
}
__device__ __forceinline__ double2 ft_twiddle_v(double2*, const double2* __restrict__ t, long long e, long long) {
    return ft_twiddle_z(t, e);
}

}  // namespace ofx
