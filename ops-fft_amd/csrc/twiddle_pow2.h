// twiddle_pow2.h - four-step twiddles W_N^e = exp(-2*pi*i*e/N) for N = 2^L (L in [3, 27]), computed
// in registers, no table.
//
// The reduction is exact integer arithmetic: m = e mod N, q = round(4m/N) (the nearest quarter
// turn) and s = m - q*N/4 in [-N/8, N/8], so |s| <= 2^24 and x = 2s/N is an exact float.  Then
// W = i^(-q) * exp(-i*pi*x) with (sin, cos)(pi*x) from sincospif on |x| <= 1/4: the quarter-turn
// rotation is a swap and sign changes, so the result carries only sincospif's own rounding
// (about 1 ulp), with no table loads.  A column kernel lane visits twiddles of unrelated
// exponents, which as table lookups cost one cache line per lane per load.
#pragma once
#include <hip/hip_runtime.h>

namespace ofx {

__device__ __forceinline__ float2 twiddle_pow2(unsigned long long e, int log2n) {
    const unsigned long long m = e & ((1ULL << log2n) - 1ULL);
    const int sh = log2n - 2;                                    // N/4 = 2^sh
    const unsigned long long q = (m + (1ULL << (sh - 1))) >> sh;  // 0..4
    const long long s = (long long)m - (long long)(q << sh);      // [-N/8, N/8]
    const float x = ldexpf((float)s, 1 - log2n);                 // 2s/N, exact
    float sn, cs;
    sincospif(x, &sn, &cs);
    // angle theta = q*pi/2 + pi*x; W = (cos theta, -sin theta)
    float ct, st;
    switch ((int)(q & 3)) {
        case 0: ct = cs; st = sn; break;
        case 1: ct = -sn; st = cs; break;
        case 2: ct = -cs; st = -sn; break;
        default: ct = sn; st = -cs; break;
    }
    return make_float2(ct, -st);
}

// complex128: the same exact reduction, FP64 sincospi (x = 2s/N is exact in double for any
// N <= 2^53), so the twiddle carries only sincospi's own rounding.
__device__ __forceinline__ double2 twiddle_pow2_z(unsigned long long e, int log2n) {
    const unsigned long long m = e & ((1ULL << log2n) - 1ULL);
    const int sh = log2n - 2;
    const unsigned long long q = (m + (1ULL << (sh - 1))) >> sh;
    const long long s = (long long)m - (long long)(q << sh);
    const double x = ldexp((double)s, 1 - log2n);
    double sn, cs;
    sincospi(x, &sn, &cs);
    double ct, st;
    switch ((int)(q & 3)) {
        case 0: ct = cs; st = sn; break;
        case 1: ct = -sn; st = cs; break;
        case 2: ct = -cs; st = -sn; break;
        default: ct = sn; st = -cs; break;
    }
    return make_double2(ct, -st);
}

// the four-step twiddle in the element type's precision
__device__ __forceinline__ float2 twiddle_pow2_v(float2*, unsigned long long e, int log2n) {
    return twiddle_pow2(e, log2n);
}
__device__ __forceinline__ double2 twiddle_pow2_v(double2*, unsigned long long e, int log2n) {
    return twiddle_pow2_z(e, log2n);
}

}  // namespace ofx
