// mfma_split.h - split-precision operands for the CDNA4 MFMA DFT kernels.
//
// Each FP32 operand value is written as a sum of 1 (BF16), 2 (FP16 hi + lo) or 3 (BF16 hi + mid +
// lo) narrow values; a product of split operands keeps the partial products whose combined order
// is at or above the working precision (FP16x2: hi*hi, hi*lo, lo*hi; BF16x3: the six with
// p + q <= 2), accumulated in FP32 smallest first.
#pragma once
#include <hip/hip_runtime.h>

#include "ofx_launch.h"

namespace ofx {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned int uintx4 __attribute__((ext_vector_type(4)));

template <int MODE>
struct SplitTraits;
template <>
struct SplitTraits<OFX_MFMA_BF16> {
    static constexpr int P = 1;
    typedef bf16x8 vec;
};
template <>
struct SplitTraits<OFX_MFMA_FP16X2> {
    static constexpr int P = 2;
    typedef half8 vec;
};
template <>
struct SplitTraits<OFX_MFMA_BF16X3> {
    static constexpr int P = 3;
    typedef bf16x8 vec;
};

// v -> P narrow parts, successive rounding of the remainder
template <int MODE>
__device__ __forceinline__ void split8(const float (&v)[8], typename SplitTraits<MODE>::vec (&out)[SplitTraits<MODE>::P]) {
    constexpr int P = SplitTraits<MODE>::P;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float rem = v[j];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if constexpr (MODE == OFX_MFMA_FP16X2) {
                const _Float16 q = (_Float16)rem;
                out[p][j] = q;
                rem -= (float)q;
            } else {
                const __bf16 q = (__bf16)rem;
                out[p][j] = q;
                rem -= (float)q;
            }
        }
    }
}

// the same split of an FP64 value (constant tables: FP64 sincospi, rounded part by part)
template <int MODE>
__device__ __forceinline__ void split8_f64(const double (&v)[8], typename SplitTraits<MODE>::vec (&out)[SplitTraits<MODE>::P]) {
    constexpr int P = SplitTraits<MODE>::P;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        double rem = v[j];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if constexpr (MODE == OFX_MFMA_FP16X2) {
                const _Float16 q = (_Float16)(float)rem;
                out[p][j] = q;
                rem -= (double)(float)q;
            } else {
                const __bf16 q = (__bf16)(float)rem;
                out[p][j] = q;
                rem -= (double)(float)q;
            }
        }
    }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Power-of-two scale that puts a transform's largest component in [2^11, 2^12): FP16 has a narrow
// exponent range, so the FP16 hi/lo operands are formed from scaled data (the scale is undone
// exactly on the output).  BF16 keeps FP32's exponent range and needs none.
template <int MODE>
__device__ __forceinline__ float fp16_scale(float wave_max_abs) {
    if constexpr (MODE != OFX_MFMA_FP16X2) {
        (void)wave_max_abs;
        return 1.f;
    } else {
        const float m = wave_max_abs;
        // (the exponent is capped so that the scale stays finite for subnormal data)
        return (m > 0.f && m < 3.0e38f) ? ldexpf(1.f, min(11 - ilogbf(m), 126)) : 1.f;
    }
}

}  // namespace ofx
