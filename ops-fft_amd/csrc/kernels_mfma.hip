// kernels_mfma.hip - C2C FFT of length 1024 as two dense 32-point DFT stages on CDNA4 MFMA
// (v_mfma_f32_32x32x16_{f16,bf16}) with split-precision operands and FP32 accumulation.
//
// Decomposition (n = n1 + 32 n2, k = 32 k1 + k2):
//   stage 1  D1[n1][k2] = sum_n2 x[n1 + 32 n2] F[n2][k2]          (MFMA, data = A operand)
//   twiddle  Z[n1][k2]  = D1[n1][k2] W_1024^(n1 k2)                (FP32 VALU)
//   stage 2  X[32 k1 + k2] = sum_n1 F[k1][n1] Z[n1][k2]            (MFMA, data = B operand)
// with F[a][b] = W_32^(ab).  Complex products use the real form: stage 1 is
// [xr | xi] (32x64) . [[Fr, Fi], [-Fi, Fr]] (64x64), stage 2 is [[Fr, -Fi], [Fi, Fr]] (64x64) .
// [Zr ; Zi] (64x32).  This is the matrix-DFT formulation of the reference's Cube-unit kernels
// (src/fft1_d/arch32/fft_b/fft1_d_b.cpp:27-150: per stage a dense 2R x 2R real DFT matrix, with
// the inter-stage twiddle on the vector unit), recast for the MFMA operand layouts.
//
// Lane layouts of v_mfma_f32_32x32x16 (wave64, lane l, r = l % 32, h = l / 32):
//   A operand: 8 values A[r][16s + 8h + j], j < 8   (k-step s covers K = 16s .. 16s+15)
//   B operand: 8 values B[16s + 8h + j][r]
//   D (16 FP32 accumulators): D[8(v/4) + 4h + v%4][r], v < 16
// Stage 1 takes its A operand straight from HBM (for a fixed j the 32 lanes of a half-wave read
// 32 consecutive complex64 values, 256 B).  Stage 1's accumulators hold, per lane, column
// k2 = r and rows n1 = 8(v/4) + 4h + v%4; stage 2 needs B[K][k2] with k2 = r, so using the lane's
// own rows as its K slots (K slot 16s + 8h + j <-> n1 of accumulator v = 8(s%2) + j) feeds stage 2
// from registers with no LDS transpose; the constant A operand of stage 2 is generated with the same
// K permutation.  Stage 2's accumulators hold X[32 k1 + r], again 256 B runs per half-wave.
//
// Split precision: each FP32 operand value is written as a sum of 1 (BF16), 2 (FP16 hi + lo) or 3
// (BF16 hi + mid + lo) narrow values; the products whose combined order is below the working
// precision are dropped (FP16x2: 3 MFMAs per tile and k-step, BF16x3: 6).  FP16 has a narrow
// exponent range, so each transform is scaled by a power of two taken from its largest component
// (max * s1 in [2^11, 2^12)), and stage 2's input by s1 * 2^-5 (|D1| <= 32 max|x|); the output is
// scaled back exactly.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ofx_launch.h"
#include "codelets.h"
#include "mfma_split.h"

namespace ofx {

constexpr int kMfmaBlock = 256;  // 4 waves, one transform per wave at a time

template <int MODE>
__device__ __forceinline__ floatx16 mfma(typename SplitTraits<MODE>::vec a, typename SplitTraits<MODE>::vec b,
                                         floatx16 c) {
    if constexpr (MODE == OFX_MFMA_FP16X2)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// acc += sum of the products a_p * b_q with p + q < P (the dropped ones are below FP32 precision)
template <int MODE>
__device__ __forceinline__ floatx16 mfma_split(const typename SplitTraits<MODE>::vec (&a)[SplitTraits<MODE>::P],
                                               const typename SplitTraits<MODE>::vec (&b)[SplitTraits<MODE>::P],
                                               floatx16 acc) {
    constexpr int P = SplitTraits<MODE>::P;
    // smallest terms first
#pragma unroll
    for (int d = P - 1; d >= 0; --d) {
#pragma unroll
        for (int p = 0; p <= d; ++p) acc = mfma<MODE>(a[p], b[d - p], acc);
    }
    return acc;
}


// Constant-operand table, generated at plan time in lane order:
//   index ((((stage * 2 + tile) * 4 + kstep) * P + part) * 64 + lane) of 16-byte vectors
// One transform: v holds its 16 lane values (loaded by the caller), the result goes to y.
template <int MODE, bool NT>
__device__ __forceinline__ void mfma_transform(const float2 (&v)[16], float2* __restrict__ y,
                                               const typename SplitTraits<MODE>::vec* cst, const float2* tw,
                                               int lane, int r, int h, float in_sign, float out_sign, float scale) {
    typedef typename SplitTraits<MODE>::vec vec;
    constexpr int P = SplitTraits<MODE>::P;
    float m = 0.f;
#pragma unroll
    for (int q = 0; q < 16; ++q) m = fmaxf(m, fmaxf(fabsf(v[q].x), fabsf(v[q].y)));
    m = wave_max(m);
    const float s1 = fp16_scale<OFX_MFMA_FP16X2>(m);  // every mode: the power-of-two scale is exact
    const float si = s1 * in_sign;

    // ---- stage 1 ------------------------------------------------------------------------------
    floatx16 acc0 = {0}, acc1 = {0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        float av[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) av[j] = (s < 2) ? v[s * 8 + j].x * s1 : v[(s - 2) * 8 + j].y * si;
        vec ap[P];
        split8<MODE>(av, ap);
        vec b0[P], b1[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            b0[p] = cst[(((0 * 2 + 0) * 4 + s) * P + p) * 64 + lane];
            b1[p] = cst[(((0 * 2 + 1) * 4 + s) * P + p) * 64 + lane];
        }
        acc0 = mfma_split<MODE>(ap, b0, acc0);
        acc1 = mfma_split<MODE>(ap, b1, acc1);
    }

    // ---- inter-stage twiddle W_1024^(n1 k2), k2 = r, and rescale by 2^-5 --------------------
    float2 z[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int n1 = 8 * (q >> 2) + 4 * h + (q & 3);
        const float2 w = tw[n1 * r];
        const float2 d = make_float2(acc0[q] * 0.03125f, acc1[q] * 0.03125f);
        z[q] = cmul(d, w);
    }

    // ---- stage 2 ------------------------------------------------------------------------------
    floatx16 out0 = {0}, out1 = {0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        float bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) bv[j] = (s < 2) ? z[s * 8 + j].x : z[(s - 2) * 8 + j].y;
        vec bp[P];
        split8<MODE>(bv, bp);
        vec a0[P], a1[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            a0[p] = cst[(((1 * 2 + 0) * 4 + s) * P + p) * 64 + lane];
            a1[p] = cst[(((1 * 2 + 1) * 4 + s) * P + p) * 64 + lane];
        }
        out0 = mfma_split<MODE>(a0, bp, out0);
        out1 = mfma_split<MODE>(a1, bp, out1);
    }

    // ---- store X[32 k1 + k2] ------------------------------------------------------------------
    const float inv = 32.f / s1 * scale;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k1 = 8 * (q >> 2) + 4 * h + (q & 3);
        const float2 o = make_float2(out0[q] * inv, out1[q] * (inv * out_sign));
        if constexpr (NT)
            st_nt(y + 32 * k1 + r, o);
        else
            y[32 * k1 + r] = o;
    }
}

// NT: streaming (non-temporal) HBM access, for working sets well beyond the last-level cache
template <bool NT>
__device__ __forceinline__ void mfma_load(float2 (&v)[16], const float2* __restrict__ x, int r, int h) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int n2 = 16 * (q >> 3) + 8 * h + (q & 7);
        v[q] = NT ? ld_nt(x + r + 32 * n2) : x[r + 32 * n2];
    }
}

// Constant-operand table, generated at plan time in lane order:
//   index ((((stage * 2 + tile) * 4 + kstep) * P + part) * 64 + lane) of 16-byte vectors
// PREFETCH: the next transform's loads are issued before the current transform's MFMA work.
template <int MODE, bool PREFETCH, bool NT>
__global__ __launch_bounds__(kMfmaBlock) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_mfma_c2c_1024(ofx_args a) {
    typedef typename SplitTraits<MODE>::vec vec;
    constexpr int P = SplitTraits<MODE>::P;
    constexpr int NCONST = 2 * 2 * 4 * P * 64;
    __shared__ vec cst[NCONST];
    __shared__ float2 tw[1024];
    {
        const vec* g = static_cast<const vec*>(a.tw);
        for (int i = threadIdx.x; i < NCONST; i += kMfmaBlock) cst[i] = g[i];
        const float2* gt = static_cast<const float2*>(a.tw_aux);
        for (int i = threadIdx.x; i < 1024; i += kMfmaBlock) tw[i] = gt[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r = lane & 31;
    const int h = lane >> 5;
    const long long stride = (long long)gridDim.x * (kMfmaBlock / 64);
    const float2* __restrict__ in = static_cast<const float2*>(a.in);
    float2* __restrict__ out = static_cast<float2*>(a.out);
    long long t = (long long)blockIdx.x * (kMfmaBlock / 64) + wave;
    if constexpr (PREFETCH) {
        float2 cur[16];
        if (t < a.batch) mfma_load<NT>(cur, in + t * a.idist, r, h);
        for (; t < a.batch; t += stride) {
            __asm__ volatile("" ::: "memory");  // keep the constant operands in LDS (see below)
            float2 nxt[16];
            const long long tn = t + stride;
            if (tn < a.batch) mfma_load<NT>(nxt, in + tn * a.idist, r, h);
            mfma_transform<MODE, NT>(cur, out + t * a.odist, cst, tw, lane, r, h, a.in_sign, a.out_sign, a.scale);
#pragma unroll
            for (int q = 0; q < 16; ++q) cur[q] = nxt[q];
        }
    } else {
        for (; t < a.batch; t += stride) {
            // Re-read the constant operands from LDS in every iteration instead of letting the
            // compiler keep all of them in VGPRs across the loop (that would leave one wave per SIMD).
            __asm__ volatile("" ::: "memory");
            float2 v[16];
            mfma_load<NT>(v, in + t * a.idist, r, h);
            mfma_transform<MODE, NT>(v, out + t * a.odist, cst, tw, lane, r, h, a.in_sign, a.out_sign, a.scale);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// constant-operand table: FP64 sincospi on the GPU, split into narrow parts
// ---------------------------------------------------------------------------------------------
// value of stage-1 B operand (K, col) and stage-2 A operand (row, K) for the 1024 = 32 x 32 plan
__device__ double mfma_const(int stage, int tile, int kslot, int lanev) {
    const int r = lanev & 31, h = lanev >> 5;
    const int K = kslot;  // 0..63
    double s, c;
    if (stage == 0) {
        // B1[K][32*tile + r]: K < 32 -> (n2 = K), K >= 32 -> (n2 = K - 32); k2 = r
        const int n2 = K & 31;
        sincospi(2.0 * (double)((n2 * r) & 31) / 32.0, &s, &c);
        const double fr = c, fi = -s;  // F = W_32^(n2 k2)
        if (tile == 0) return (K < 32) ? fr : -fi;
        return (K < 32) ? fi : fr;
    }
    // stage 2: A2[32*tile + r][K] with K slot -> n1 through the accumulator permutation
    const int kk = K & 31;                 // slot within the re (or im) half
    const int sh = kk >> 4;                // k-step within the half
    const int hh = (kk >> 3) & 1;          // which half-wave supplies this slot
    const int j = kk & 7;
    const int vacc = 8 * sh + j;           // accumulator index of that lane
    const int n1 = 8 * (vacc >> 2) + 4 * hh + (vacc & 3);
    (void)h;
    const int k1 = r;
    sincospi(2.0 * (double)((k1 * n1) & 31) / 32.0, &s, &c);
    const double fr = c, fi = -s;
    if (tile == 0) return (K < 32) ? fr : -fi;
    return (K < 32) ? fi : fr;
}

template <int MODE>
__global__ void k_gen_mfma_consts(void* out) {
    typedef typename SplitTraits<MODE>::vec vec;
    constexpr int P = SplitTraits<MODE>::P;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (stage, tile, kstep, lane)
    if (idx >= 2 * 2 * 4 * 64) return;
    const int lane = idx & 63;
    const int ks = (idx >> 6) & 3;
    const int tile = (idx >> 8) & 1;
    const int stage = idx >> 9;
    const int h = lane >> 5;
    vec parts[P];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int kslot = 16 * ks + 8 * h + j;
        double rem = mfma_const(stage, tile, kslot, lane);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if constexpr (MODE == OFX_MFMA_FP16X2) {
                const _Float16 q = (_Float16)(float)rem;
                parts[p][j] = q;
                rem -= (double)(float)q;
            } else {
                const __bf16 q = (__bf16)(float)rem;
                parts[p][j] = q;
                rem -= (double)(float)q;
            }
        }
    }
    vec* o = static_cast<vec*>(out);
#pragma unroll
    for (int p = 0; p < P; ++p) o[(((stage * 2 + tile) * 4 + ks) * P + p) * 64 + lane] = parts[p];
}

}  // namespace ofx

using namespace ofx;

extern "C" long long ofx_mfma_const_bytes(int mode) {
    const int P = mode == OFX_MFMA_BF16 ? 1 : (mode == OFX_MFMA_FP16X2 ? 2 : (mode == OFX_MFMA_BF16X3 ? 3 : 0));
    return (long long)2 * 2 * 4 * P * 64 * 16;
}

extern "C" int ofx_gen_mfma_consts(int mode, void* out, hipStream_t stream) {
    switch (mode) {
        case OFX_MFMA_BF16: hipLaunchKernelGGL(k_gen_mfma_consts<OFX_MFMA_BF16>, dim3(4), dim3(256), 0, stream, out); break;
        case OFX_MFMA_FP16X2: hipLaunchKernelGGL(k_gen_mfma_consts<OFX_MFMA_FP16X2>, dim3(4), dim3(256), 0, stream, out); break;
        case OFX_MFMA_BF16X3: hipLaunchKernelGGL(k_gen_mfma_consts<OFX_MFMA_BF16X3>, dim3(4), dim3(256), 0, stream, out); break;
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

extern "C" int ofx_launch_mfma_c2c(int mode, int n, const ofx_args* a, int blocks, hipStream_t stream) {
    if (a == nullptr || n != 1024) return (int)hipErrorInvalidValue;
    if (a->batch <= 0) return 0;
    const long long need = (a->batch + 3) / 4;
    const bool prefetch = blocks < 0;
    const int nb = blocks < 0 ? -blocks : blocks;
    // default grid (measured on MI355X): 2048 workgroups for the 1- and 2-part modes, 512 for the
    // 3-part mode whose constant table takes 48 KiB of LDS per workgroup
    const int dflt = (mode == OFX_MFMA_BF16X3) ? 512 : 2048;
    const unsigned g = (unsigned)std::min<long long>(need, nb > 0 ? nb : dflt);
    // streaming HBM access when input and output together exceed the 256 MB last-level cache by
    // half again (the threshold of the FP32 kernels' streaming variants)
    const bool nt = a->batch * 1024LL * 16 >= (384LL << 20);
#define OFX_MFMA_LAUNCH(M)                                                                                   \
    do {                                                                                                     \
        if (prefetch && nt)                                                                                  \
            hipLaunchKernelGGL((k_mfma_c2c_1024<M, true, true>), dim3(g), dim3(kMfmaBlock), 0, stream, *a);  \
        else if (prefetch)                                                                                   \
            hipLaunchKernelGGL((k_mfma_c2c_1024<M, true, false>), dim3(g), dim3(kMfmaBlock), 0, stream, *a); \
        else if (nt)                                                                                         \
            hipLaunchKernelGGL((k_mfma_c2c_1024<M, false, true>), dim3(g), dim3(kMfmaBlock), 0, stream, *a); \
        else                                                                                                 \
            hipLaunchKernelGGL((k_mfma_c2c_1024<M, false, false>), dim3(g), dim3(kMfmaBlock), 0, stream, *a); \
    } while (0)
    switch (mode) {
        case OFX_MFMA_BF16: OFX_MFMA_LAUNCH(OFX_MFMA_BF16); break;
        case OFX_MFMA_FP16X2: OFX_MFMA_LAUNCH(OFX_MFMA_FP16X2); break;
        case OFX_MFMA_BF16X3: OFX_MFMA_LAUNCH(OFX_MFMA_BF16X3); break;
        default: return (int)hipErrorInvalidValue;
    }
#undef OFX_MFMA_LAUNCH
    return (int)hipGetLastError();
}
