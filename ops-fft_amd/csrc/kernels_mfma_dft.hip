// kernels_mfma_dft.hip - C2C FFTs of length n = R1 * R2 (R1, R2 in {16, 32, 64}) as two dense DFT
// stages on CDNA4 MFMA (v_mfma_f32_16x16x32_{f16,bf16}), split-precision operands, FP32 accumulate.
//
// These are the 16-, 32- and 64-point matrix-DFT stages of the reference's Cube kernels: its
// radix pairs {16,16} for 256, {16,32} for 512, {16,64} for 1024, {32,64} for 2048 and {64,64} for
// 4096 (src/fft1_d/arch32/fft_b/fft1_d_b.cpp:27-48), each stage a dense 2R x 2R real DFT matrix
// applied to re/im-separated data with the inter-stage twiddle on the vector unit
// (fft_b/fft1_d_b.cpp:50-150), recast for the MFMA operand layouts.
//
// Decomposition (n = n1 + R1 n2, k = R2 k1 + k2; F_R[a][b] = W_R^(ab)):
//   stage 1  D[n1][k2] = sum_n2 x[n1 + R1 n2] F_R2[n2][k2]      (data = A operand)
//   twiddle  Z[n1][k2] = D[n1][k2] W_n^(n1 k2)                   (FP32 VALU)
//   stage 2  X[R2 k1 + k2] = sum_n1 F_R1[k1][n1] Z[n1][k2]       (data = B operand)
//
// Lane layout of v_mfma_f32_16x16x32 (wave64, lane l, r = l % 16, g = l / 16): the A operand holds
// row r, the B operand column r, each with 8 K values of lane group g; the four accumulators hold
// D[4g + v][r], v < 4.  K slots are a free permutation as long as A and B use the same one, so
// slot (g, j) stands for (j < 4 ? real : imaginary part) of input index 16s + 4g + (j & 3) in
// k-step s.  Then:
//   - stage 1 loads its A operand straight from HBM: lane (r, g) reads x[(16mt + r) + R1 (16s + 4g
//     + q)], q < 4, complex pairs (16 lanes read 128 contiguous bytes);
//   - stage 1's accumulators give the lane D[16mt + 4g + v][16c + r] for every k2 block c, real and
//     imaginary parts: exactly the B-operand values (slot (g, j) <-> n1 = 16mt + 4g + (j & 3)) that
//     k-step mt of stage 2 needs for column block c.  Stage 2 is therefore accumulated M-tile by
//     M-tile of stage 1, with no LDS exchange, and the stage-2 constant operand is generated with
//     the same slot map;
//   - stage 2's accumulators give X[R2 (16b + 4g + v) + 16c + r]: 16 lanes store 128 contiguous
//     bytes.
// Real form: a complex product P = F Y becomes [Pr | Pi] with Pr = Fr Yr - Fi Yi and Pi = Fi Yr +
// Fr Yi.  Per lane, the constant vector of the real-part output tile is [Fr(q) | -Fi(q)] over its
// four real and four imaginary slots; the imaginary-part tile is [Fi(q) | Fr(q)], derived from the
// stored one by swapping halves and flipping two sign bits, so only the real-part tiles are
// stored (half the LDS).
//
// Ranges: for FP16 operands each transform is scaled by a power of two (largest component in
// [2^11, 2^12)); stage 1 grows values by at most 2 R2, so its output is scaled by 1 / (2 R2)
// before the FP16 split of stage 2; both scales are undone exactly on the output.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "codelets.h"
#include "mfma_split.h"
#include "ofx_launch.h"

namespace ofx {

constexpr int kDftBlock = 256;  // 4 waves, one transform per wave at a time

template <int MODE>
__device__ __forceinline__ floatx4 mfma16(typename SplitTraits<MODE>::vec a, typename SplitTraits<MODE>::vec b,
                                          floatx4 c) {
    if constexpr (MODE == OFX_MFMA_FP16X2)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// acc += the partial products a_p * b_q with p + q < P, smallest first
template <int MODE>
__device__ __forceinline__ floatx4 mfma16_split(const typename SplitTraits<MODE>::vec (&a)[SplitTraits<MODE>::P],
                                                const typename SplitTraits<MODE>::vec (&b)[SplitTraits<MODE>::P],
                                                floatx4 acc) {
    constexpr int P = SplitTraits<MODE>::P;
#pragma unroll
    for (int d = P - 1; d >= 0; --d) {
#pragma unroll
        for (int p = 0; p <= d; ++p) acc = mfma16<MODE>(a[p], b[d - p], acc);
    }
    return acc;
}

// [x0..x3 | y0..y3] -> [-y0..-y3 | x0..x3] (imaginary-part tile from the real-part tile); the
// sign of a narrow value is its top bit
template <class vec>
__device__ __forceinline__ vec im_tile(vec re) {
    const uintx4 u = __builtin_bit_cast(uintx4, re);
    const uintx4 o = {u[2] ^ 0x80008000u, u[3] ^ 0x80008000u, u[0], u[1]};
    return __builtin_bit_cast(vec, o);
}

template <int R1_, int R2_>
struct DftShape {
    static constexpr int R1 = R1_, R2 = R2_, N = R1 * R2;
    static constexpr int S1 = R2 / 16;  // stage-1 k-steps = stage-1 column blocks = stage-2 column blocks
    static constexpr int S2 = R1 / 16;  // stage-1 row tiles = stage-2 k-steps = stage-2 row blocks
    static constexpr bool TW2 = N > 2048;  // two-level twiddle table (2 x 64 entries) instead of N
    static constexpr int NTW = TW2 ? 128 : N;
    static_assert(R1 % 16 == 0 && R2 % 16 == 0 && R1 <= 64 && R2 <= 64, "radices 16, 32 or 64");
};

// constant vectors per part: stage 1 (column block c, k-step s), then stage 2 (row block b, k-step s)
template <class SH, int P>
constexpr int dft_nconst() {
    return (SH::S1 * SH::S1 + SH::S2 * SH::S2) * P * 64;
}

template <bool NT>
__device__ __forceinline__ float2 ldx(const float2* p) {
    if constexpr (NT)
        return ld_nt(p);
    else
        return *p;
}

// 2D store (OFX_FLAG_MFMA_2D): the input is an image of R2 rows and R1 columns, stage 1 transformed
// its columns and stage 2 its rows, so accumulator (b, c, q) is output element (k2, k1) = (16 c + r,
// 16 b + 4 g + q) of the image, at R1 k2 + k1.  A lane's four values q are then 32 contiguous bytes
// of one row: with a 16-byte aligned output (v16) they go out as two 16-byte stores, and four lanes
// cover a 128-byte row segment.
template <class SH, bool NT>
__device__ __forceinline__ void dft_store_2d(const floatx4 (&acc2)[SH::S2][SH::S1][2], float2* __restrict__ y,
                                             int r, int g, float inv, float invi, bool v16) {
#pragma unroll
    for (int b = 0; b < SH::S2; ++b)
#pragma unroll
        for (int c = 0; c < SH::S1; ++c) {
            float2* row = y + SH::R1 * (16 * c + r) + 16 * b + 4 * g;
            if (v16) {
#pragma unroll
                for (int q = 0; q < 4; q += 2) {
                    const float4 val = make_float4(acc2[b][c][0][q] * inv, acc2[b][c][1][q] * invi,
                                                   acc2[b][c][0][q + 1] * inv, acc2[b][c][1][q + 1] * invi);
                    if constexpr (NT)
                        st_nt(reinterpret_cast<float4*>(row + q), val);
                    else
                        *reinterpret_cast<float4*>(row + q) = val;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float2 val = make_float2(acc2[b][c][0][q] * inv, acc2[b][c][1][q] * invi);
                    if constexpr (NT)
                        st_nt(row + q, val);
                    else
                        row[q] = val;
                }
            }
        }
}

template <int MODE, class SH, bool NT, bool T2D>
__device__ __forceinline__ void dft_transform(const float2 (&v)[SH::S2][SH::S1][4], float2* __restrict__ y,
                                              const typename SplitTraits<MODE>::vec* cst, const float2* tw, int lane,
                                              int r, int g, float in_sign, float out_sign, float scale, bool v16) {
    typedef typename SplitTraits<MODE>::vec vec;
    constexpr int P = SplitTraits<MODE>::P;
    constexpr int S1 = SH::S1, S2 = SH::S2, R2 = SH::R2;
    constexpr int NC1 = S1 * S1 * P * 64;
    float s1 = 1.f;
    if constexpr (MODE == OFX_MFMA_FP16X2) {
        float m = 0.f;
#pragma unroll
        for (int a = 0; a < S2; ++a)
#pragma unroll
            for (int b = 0; b < S1; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) m = fmaxf(m, fmaxf(fabsf(v[a][b][q].x), fabsf(v[a][b][q].y)));
        s1 = fp16_scale<MODE>(wave_max(m));
    }
    const float si = s1 * in_sign;
    constexpr float zs = 1.f / (2 * R2);

    floatx4 acc2[S2][S1][2];
#pragma unroll
    for (int b = 0; b < S2; ++b)
#pragma unroll
        for (int c = 0; c < S1; ++c) acc2[b][c][0] = acc2[b][c][1] = floatx4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
    for (int mt = 0; mt < S2; ++mt) {
        // the constant operands are re-read from LDS for every row tile instead of being held in
        // VGPRs across tiles (for 64-point stages they would take 128 VGPRs)
        __asm__ volatile("" ::: "memory");
        // ---- stage 1, rows n1 = 16 mt + (0..15) ---------------------------------------------
        floatx4 acc1[S1][2];
#pragma unroll
        for (int c = 0; c < S1; ++c) acc1[c][0] = acc1[c][1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < S1; ++s) {
            float av[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = v[mt][s][q].x * s1;
                av[4 + q] = v[mt][s][q].y * si;
            }
            vec ap[P];
            split8<MODE>(av, ap);
#pragma unroll
            for (int c = 0; c < S1; ++c) {
                vec bre[P], bim[P];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    bre[p] = cst[((c * S1 + s) * P + p) * 64 + lane];
                    bim[p] = im_tile(bre[p]);
                }
                acc1[c][0] = mfma16_split<MODE>(ap, bre, acc1[c][0]);
                acc1[c][1] = mfma16_split<MODE>(ap, bim, acc1[c][1]);
            }
        }
        // ---- twiddle W_n^(n1 k2), then k-step mt of stage 2 ---------------------------------
#pragma unroll
        for (int c = 0; c < S1; ++c) {
            float bv[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = (16 * mt + 4 * g + q) * (16 * c + r);
                float2 w;
                if constexpr (SH::TW2)
                    w = cmul(tw[e >> 6], tw[64 + (e & 63)]);
                else
                    w = tw[e];
                const float2 z = cmul(make_float2(acc1[c][0][q] * zs, acc1[c][1][q] * zs), w);
                bv[q] = z.x;
                bv[4 + q] = z.y;
            }
            vec bp[P];
            split8<MODE>(bv, bp);
#pragma unroll
            for (int b = 0; b < S2; ++b) {
                vec are[P], aim[P];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    are[p] = cst[NC1 + ((b * S2 + mt) * P + p) * 64 + lane];
                    aim[p] = im_tile(are[p]);
                }
                acc2[b][c][0] = mfma16_split<MODE>(are, bp, acc2[b][c][0]);
                acc2[b][c][1] = mfma16_split<MODE>(aim, bp, acc2[b][c][1]);
            }
        }
    }

    // ---- X[R2 k1 + k2], k1 = 16 b + 4 g + q, k2 = 16 c + r (2D: dft_store_2d) ----------------
    const float inv = (2.f * R2) / s1 * scale;
    const float invi = inv * out_sign;
    if constexpr (T2D) {
        dft_store_2d<SH, NT>(acc2, y, r, g, inv, invi, v16);
        return;
    }
#pragma unroll
    for (int b = 0; b < S2; ++b)
#pragma unroll
        for (int c = 0; c < S1; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float2* o = y + R2 * (16 * b + 4 * g + q) + 16 * c + r;
                const float2 val = make_float2(acc2[b][c][0][q] * inv, acc2[b][c][1][q] * invi);
                if constexpr (NT)
                    st_nt(o, val);
                else
                    *o = val;
            }
}

// Tiled form: the transform is read one stage-1 row tile (16 rows n1, all R2 columns: R2/4
// complex values per lane) at a time, the next tile's loads issued before the current tile's MFMA
// work, so only one tile of input is held in registers besides the stage-2 accumulators.  FP16
// operands then take a power-of-two scale per tile; stage 2 keeps one running scale (the smallest
// so far): a tile with a smaller scale rescales the accumulators, one with a larger scale has its
// stage-2 input scaled down to the running one.  Every scale is a power of two, so this is exact.
template <class SH, bool NT>
__device__ __forceinline__ void dft_load_tile(float2 (&v)[SH::S1][4], const float2* __restrict__ x, int mt, int r,
                                              int g) {
#pragma unroll
    for (int s = 0; s < SH::S1; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[s][q] = ldx<NT>(x + (16 * mt + r) + SH::R1 * (16 * s + 4 * g + q));
}

template <int MODE, class SH, bool NT, bool T2D>
__device__ __forceinline__ void dft_transform_tiled(const float2* __restrict__ x, float2* __restrict__ y,
                                                    const typename SplitTraits<MODE>::vec* cst, const float2* tw,
                                                    int lane, int r, int g, float in_sign, float out_sign,
                                                    float scale, bool v16) {
    typedef typename SplitTraits<MODE>::vec vec;
    constexpr int P = SplitTraits<MODE>::P;
    constexpr int S1 = SH::S1, S2 = SH::S2, R2 = SH::R2;
    constexpr int NC1 = S1 * S1 * P * 64;
    constexpr float zs = 1.f / (2 * R2);

    floatx4 acc2[S2][S1][2];
#pragma unroll
    for (int b = 0; b < S2; ++b)
#pragma unroll
        for (int c = 0; c < S1; ++c) acc2[b][c][0] = acc2[b][c][1] = floatx4{0.f, 0.f, 0.f, 0.f};
    float srun = 1.f;  // running stage-2 scale (FP16 operands)

    float2 cur[S1][4];
    dft_load_tile<SH, NT>(cur, x, 0, r, g);
#pragma unroll
    for (int mt = 0; mt < S2; ++mt) {
        __asm__ volatile("" ::: "memory");  // constants re-read per tile; loads stay one tile ahead
        float2 nxt[S1][4];
        if (mt + 1 < S2) dft_load_tile<SH, NT>(nxt, x, mt + 1, r, g);
        float st = 1.f;
        if constexpr (MODE == OFX_MFMA_FP16X2) {
            float m = 0.f;
#pragma unroll
            for (int s = 0; s < S1; ++s)
#pragma unroll
                for (int q = 0; q < 4; ++q) m = fmaxf(m, fmaxf(fabsf(cur[s][q].x), fabsf(cur[s][q].y)));
            st = fp16_scale<MODE>(wave_max(m));
        }
        const float si = st * in_sign;
        floatx4 acc1[S1][2];
#pragma unroll
        for (int c = 0; c < S1; ++c) acc1[c][0] = acc1[c][1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < S1; ++s) {
            float av[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = cur[s][q].x * st;
                av[4 + q] = cur[s][q].y * si;
            }
            vec ap[P];
            split8<MODE>(av, ap);
#pragma unroll
            for (int c = 0; c < S1; ++c) {
                vec bre[P], bim[P];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    bre[p] = cst[((c * S1 + s) * P + p) * 64 + lane];
                    bim[p] = im_tile(bre[p]);
                }
                acc1[c][0] = mfma16_split<MODE>(ap, bre, acc1[c][0]);
                acc1[c][1] = mfma16_split<MODE>(ap, bim, acc1[c][1]);
            }
        }
        float zt = zs;
        if constexpr (MODE == OFX_MFMA_FP16X2) {
            if (mt == 0) {
                srun = st;
            } else if (st < srun) {  // wave-uniform
                const float ratio = st / srun;
#pragma unroll
                for (int b = 0; b < S2; ++b)
#pragma unroll
                    for (int c = 0; c < S1; ++c) {
                        acc2[b][c][0] *= ratio;
                        acc2[b][c][1] *= ratio;
                    }
                srun = st;
            }
            zt = zs * (srun / st);
        }
#pragma unroll
        for (int c = 0; c < S1; ++c) {
            float bv[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = (16 * mt + 4 * g + q) * (16 * c + r);
                float2 w;
                if constexpr (SH::TW2)
                    w = cmul(tw[e >> 6], tw[64 + (e & 63)]);
                else
                    w = tw[e];
                const float2 z = cmul(make_float2(acc1[c][0][q] * zt, acc1[c][1][q] * zt), w);
                bv[q] = z.x;
                bv[4 + q] = z.y;
            }
            vec bp[P];
            split8<MODE>(bv, bp);
#pragma unroll
            for (int b = 0; b < S2; ++b) {
                vec are[P], aim[P];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    are[p] = cst[NC1 + ((b * S2 + mt) * P + p) * 64 + lane];
                    aim[p] = im_tile(are[p]);
                }
                acc2[b][c][0] = mfma16_split<MODE>(are, bp, acc2[b][c][0]);
                acc2[b][c][1] = mfma16_split<MODE>(aim, bp, acc2[b][c][1]);
            }
        }
        if (mt + 1 < S2) {
#pragma unroll
            for (int s = 0; s < S1; ++s)
#pragma unroll
                for (int q = 0; q < 4; ++q) cur[s][q] = nxt[s][q];
        }
    }

    const float inv = (2.f * R2) / srun * scale;
    const float invi = inv * out_sign;
    if constexpr (T2D) {
        dft_store_2d<SH, NT>(acc2, y, r, g, inv, invi, v16);
        return;
    }
#pragma unroll
    for (int b = 0; b < S2; ++b)
#pragma unroll
        for (int c = 0; c < S1; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float2* o = y + R2 * (16 * b + 4 * g + q) + 16 * c + r;
                const float2 val = make_float2(acc2[b][c][0][q] * inv, acc2[b][c][1][q] * invi);
                if constexpr (NT)
                    st_nt(o, val);
                else
                    *o = val;
            }
}

template <class SH, bool NT>
__device__ __forceinline__ void dft_load(float2 (&v)[SH::S2][SH::S1][4], const float2* __restrict__ x, int r, int g) {
#pragma unroll
    for (int mt = 0; mt < SH::S2; ++mt)
#pragma unroll
        for (int s = 0; s < SH::S1; ++s)
#pragma unroll
            for (int q = 0; q < 4; ++q) v[mt][s][q] = ldx<NT>(x + (16 * mt + r) + SH::R1 * (16 * s + 4 * g + q));
}

// args: tw = constant table (ofx_gen_mfma_dft_consts), tw_aux = W_n^k for k < n (ofx_gen_roots)
template <int MODE, int R1, int R2, bool NT, bool TILED, bool T2D>
__global__ __launch_bounds__(kDftBlock) void k_mfma_dft(ofx_args a) {
    typedef DftShape<R1, R2> SH;
    typedef typename SplitTraits<MODE>::vec vec;
    constexpr int P = SplitTraits<MODE>::P;
    constexpr int NC = dft_nconst<SH, P>();
    __shared__ vec cst[NC];
    __shared__ float2 tw[SH::NTW];
    {
        const vec* gc = static_cast<const vec*>(a.tw);
        for (int i = threadIdx.x; i < NC; i += kDftBlock) cst[i] = gc[i];
        const float2* gt = static_cast<const float2*>(a.tw_aux);
        for (int i = threadIdx.x; i < SH::NTW; i += kDftBlock) {
            if constexpr (SH::TW2)
                tw[i] = i < 64 ? gt[64 * i] : gt[i - 64];
            else
                tw[i] = gt[i];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r = lane & 15;
    const int g = lane >> 4;
    const float2* __restrict__ in = static_cast<const float2*>(a.in);
    float2* __restrict__ out = static_cast<float2*>(a.out);
    // plain launch: the grid strides over the batch; tile-list launch (uneven batches): workgroup b
    // runs tiles b, b + grid, ... (`batch` tiles), each of `count` contiguous transforms at element
    // offset `offset`, one wave per transform, so the constant tables are loaded once per
    // workgroup, not once per tile
    const bool tiled_list = a.tiles != nullptr;
    // T2D (OFX_FLAG_MFMA_2D): a compiled variant of its own, so the 1D kernels keep their registers;
    // 16-byte stores of the 2D layout need a 16-byte aligned output (transform offsets are even)
    const bool v16 = T2D && (reinterpret_cast<unsigned long long>(a.out) & 15) == 0 && (a.odist & 1) == 0;
    const long long ntile = tiled_list ? a.batch : 1;
    for (long long ti = tiled_list ? blockIdx.x : 0; ti < ntile; ti += tiled_list ? gridDim.x : 1) {
        const float2* __restrict__ tin = in;
        float2* __restrict__ tout = out;
        long long t = (long long)blockIdx.x * (kDftBlock / 64) + wave, stride = (long long)gridDim.x * (kDftBlock / 64);
        long long end = a.batch;
        if (tiled_list) {
            const ofx_batch_tile tile = static_cast<const ofx_batch_tile*>(a.tiles)[ti];
            tin += tile.offset;
            tout += tile.offset;
            t = wave;
            stride = kDftBlock / 64;
            end = tile.count;
        }
        for (; t < end; t += stride) {
            // constants are re-read from LDS for every transform rather than held in VGPRs
            __asm__ volatile("" ::: "memory");
            if constexpr (TILED) {
                dft_transform_tiled<MODE, SH, NT, T2D>(tin + t * a.idist, tout + t * a.odist, cst, tw, lane, r, g,
                                                       a.in_sign, a.out_sign, a.scale, v16);
            } else {
                float2 v[SH::S2][SH::S1][4];
                dft_load<SH, NT>(v, tin + t * a.idist, r, g);
                dft_transform<MODE, SH, NT, T2D>(v, tout + t * a.odist, cst, tw, lane, r, g, a.in_sign, a.out_sign,
                                                 a.scale, v16);
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Dense single-stage DFT for n = 16, 32, 64 (the reference's dense small-n matrix DFT,
// src/fft1_d/arch32/dft/fft1_d_dft.cpp:23-77): Y = X F_n for 16 transforms at a time per wave, the
// transforms as the MFMA's M dimension.  Lane (r, g) holds the A-operand row of transform r
// (inputs 16 s + 4 g + q, four contiguous complex values per k-step), the B operands are the
// stage-1 constant tiles of F_n (the same table as a two-stage kernel's stage 1 with R2 = n), and
// accumulator v of column block c is output 16 c + r of transform 4 g + v (16 lanes store 128
// contiguous bytes).  FP16 operands take a power-of-two scale per transform (the four lanes of a
// row agree on it through two lane exchanges), undone exactly on the output.
// ---------------------------------------------------------------------------------------------
template <int MODE, int N, bool NT>
__global__ __launch_bounds__(kDftBlock) void k_mfma_dense(ofx_args a) {
    typedef typename SplitTraits<MODE>::vec vec;
    constexpr int P = SplitTraits<MODE>::P;
    constexpr int S = N / 16;
    constexpr int NC = S * S * P * 64;
    static_assert(N % 16 == 0 && N <= 64, "dense stages of 16, 32 or 64 points");
    __shared__ vec cst[NC];
    {
        const vec* gc = static_cast<const vec*>(a.tw);
        for (int i = threadIdx.x; i < NC; i += kDftBlock) cst[i] = gc[i];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r = lane & 15;
    const int g = lane >> 4;
    const float2* __restrict__ in = static_cast<const float2*>(a.in);
    float2* __restrict__ out = static_cast<float2*>(a.out);
    // 16-byte loads of input pairs when every transform starts 16-byte aligned
    const bool v16 = (reinterpret_cast<unsigned long long>(a.in) & 15) == 0 && (a.idist & 1) == 0;
    const long long groups = (a.batch + 15) / 16;
    for (long long grp = (long long)blockIdx.x * (kDftBlock / 64) + wave; grp < groups;
         grp += (long long)gridDim.x * (kDftBlock / 64)) {
        __asm__ volatile("" ::: "memory");  // constants re-read from LDS for every group
        const long long t0 = grp * 16;
        const bool okr = t0 + r < a.batch;
        float2 v[S][4];
        const float2* src = in + (okr ? (t0 + r) : 0) * a.idist + 4 * g;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!okr) {
#pragma unroll
                for (int q = 0; q < 4; ++q) v[s][q] = make_float2(0.f, 0.f);
            } else if (v16) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const float4* p4 = reinterpret_cast<const float4*>(src + 16 * s + 2 * h);
                    float4 w;
                    if constexpr (NT)
                        w = ld_nt(p4);
                    else
                        w = *p4;
                    v[s][2 * h] = make_float2(w.x, w.y);
                    v[s][2 * h + 1] = make_float2(w.z, w.w);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) v[s][q] = ldx<NT>(src + 16 * s + q);
            }
        }
        float sc = 1.f;
        if constexpr (MODE == OFX_MFMA_FP16X2) {
            float m = 0.f;
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
                for (int q = 0; q < 4; ++q) m = fmaxf(m, fmaxf(fabsf(v[s][q].x), fabsf(v[s][q].y)));
            m = fmaxf(m, __shfl_xor(m, 16, 64));
            m = fmaxf(m, __shfl_xor(m, 32, 64));
            sc = fp16_scale<MODE>(m);
        }
        const float si = sc * a.in_sign;
        floatx4 acc[S][2];
#pragma unroll
        for (int c = 0; c < S; ++c) acc[c][0] = acc[c][1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < S; ++s) {
            float av[8];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = v[s][q].x * sc;
                av[4 + q] = v[s][q].y * si;
            }
            vec ap[P];
            split8<MODE>(av, ap);
#pragma unroll
            for (int c = 0; c < S; ++c) {
                vec bre[P], bim[P];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    bre[p] = cst[((c * S + s) * P + p) * 64 + lane];
                    bim[p] = im_tile(bre[p]);
                }
                acc[c][0] = mfma16_split<MODE>(ap, bre, acc[c][0]);
                acc[c][1] = mfma16_split<MODE>(ap, bim, acc[c][1]);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long long t = t0 + 4 * g + q;
            // the scale of transform 4 g + q is held by lane 4 g + q (row r = 4 g + q)
            const float st = __shfl(sc, 4 * g + q, 64);
            const float inv = a.scale / st;
            const float invi = inv * a.out_sign;
            if (t < a.batch) {
                float2* o = out + t * a.odist + r;
#pragma unroll
                for (int c = 0; c < S; ++c) {
                    const float2 val = make_float2(acc[c][0][q] * inv, acc[c][1][q] * invi);
                    if constexpr (NT)
                        st_nt(o + 16 * c, val);
                    else
                        o[16 * c] = val;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// constant table: FP64 sincospi on the GPU, split into narrow parts
//   stage 1 (R = R2): entry ((c * S1 + s) * P + p) * 64 + lane, column k2 = 16c + r, slot j of
//     input n2 = 16s + 4g + (j & 3): j < 4 -> cos(2 pi n2 k2 / R2), j >= 4 -> sin(...)
//   stage 2 (R = R1): entry NC1 + ((b * S2 + s) * P + p) * 64 + lane, row k1 = 16b + r, slot j of
//     input n1 = 16s + 4g + (j & 3): the same with (k1 n1 mod R1)
// (the real-part tile [Fr | -Fi] with F = W_R^(ab) = cos - i sin)
// ---------------------------------------------------------------------------------------------
template <int MODE>
__global__ void k_gen_dft_consts(void* out, int R1, int R2) {
    typedef typename SplitTraits<MODE>::vec vec;
    constexpr int P = SplitTraits<MODE>::P;
    const int S1 = R2 / 16, S2 = R1 / 16;
    const int n1 = S1 * S1 * 64, n2 = S2 * S2 * 64;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (stage, block, k-step, lane)
    if (idx >= n1 + n2) return;
    const bool st2 = idx >= n1;
    const int e = st2 ? idx - n1 : idx;
    const int S = st2 ? S2 : S1;
    const int R = st2 ? R1 : R2;
    const int lane = e & 63;
    const int s = (e >> 6) % S;
    const int blk = (e >> 6) / S;
    const int r = lane & 15, g = lane >> 4;
    const int ab = 16 * blk + r;
    double vals[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int kk = 16 * s + 4 * g + (j & 3);
        double sn, cs;
        sincospi(2.0 * (double)((ab * kk) % R) / (double)R, &sn, &cs);
        vals[j] = (j < 4) ? cs : sn;
    }
    vec parts[P];
    split8_f64<MODE>(vals, parts);
    vec* o = static_cast<vec*>(out) + (st2 ? n1 * P : 0);
#pragma unroll
    for (int p = 0; p < P; ++p) o[((blk * S + s) * P + p) * 64 + lane] = parts[p];
}

// ---------------------------------------------------------------------------------------------
// registry: id 0 is the 32 x 32 kernel on v_mfma_f32_32x32x16 (kernels_mfma.hip), ids 1.. the
// two-stage kernels above
// ---------------------------------------------------------------------------------------------
struct DftEntry {
    int n, r1, r2, tiled;
};
constexpr DftEntry kDft[] = {
    {1024, 32, 32, 0},  // 32x32x16 kernel
    {256, 16, 16, 0},   {512, 16, 32, 0},   {512, 32, 16, 0},   {1024, 32, 32, 0},  {1024, 16, 64, 0},
    {1024, 64, 16, 0},  {2048, 32, 64, 0},  {2048, 64, 32, 0},  {4096, 64, 64, 0},  {1024, 32, 32, 1},
    {1024, 64, 16, 1},  {2048, 32, 64, 1},  {2048, 64, 32, 1},  {4096, 64, 64, 1},
    // dense single stage (r1 = 1): k_mfma_dense
    {16, 1, 16, 0},     {32, 1, 32, 0},     {64, 1, 64, 0},
};
constexpr int kDftCount = (int)(sizeof(kDft) / sizeof(kDft[0]));

int parts_of(int mode) {
    return mode == OFX_MFMA_BF16 ? 1 : mode == OFX_MFMA_FP16X2 ? 2 : mode == OFX_MFMA_BF16X3 ? 3 : 0;
}

typedef void (*dft_kernel_t)(ofx_args);

template <int MODE>
dft_kernel_t dft_kernel_for(int id, bool nt, bool t2d) {
#define OFX_DFT_CASE(I, A, B, T)                                                                      \
    case I:                                                                                           \
        if (t2d) return nt ? k_mfma_dft<MODE, A, B, true, T, true> : k_mfma_dft<MODE, A, B, false, T, true>; \
        return nt ? k_mfma_dft<MODE, A, B, true, T, false> : k_mfma_dft<MODE, A, B, false, T, false>
    switch (id) {
        OFX_DFT_CASE(1, 16, 16, false);
        OFX_DFT_CASE(2, 16, 32, false);
        OFX_DFT_CASE(3, 32, 16, false);
        OFX_DFT_CASE(4, 32, 32, false);
        OFX_DFT_CASE(5, 16, 64, false);
        OFX_DFT_CASE(6, 64, 16, false);
        OFX_DFT_CASE(7, 32, 64, false);
        OFX_DFT_CASE(8, 64, 32, false);
        OFX_DFT_CASE(9, 64, 64, false);
        OFX_DFT_CASE(10, 32, 32, true);
        OFX_DFT_CASE(11, 64, 16, true);
        OFX_DFT_CASE(12, 32, 64, true);
        OFX_DFT_CASE(13, 64, 32, true);
        OFX_DFT_CASE(14, 64, 64, true);
        case 15: return nt ? k_mfma_dense<MODE, 16, true> : k_mfma_dense<MODE, 16, false>;
        case 16: return nt ? k_mfma_dense<MODE, 32, true> : k_mfma_dense<MODE, 32, false>;
        case 17: return nt ? k_mfma_dense<MODE, 64, true> : k_mfma_dense<MODE, 64, false>;
        default: return nullptr;
    }
#undef OFX_DFT_CASE
}

}  // namespace ofx

using namespace ofx;

extern "C" int ofx_mfma_kernel_count(void) { return kDftCount; }

extern "C" int ofx_mfma_kernel_info_get(int id, ofx_mfma_info* info) {
    if (info == nullptr || id < 0 || id >= kDftCount) return (int)hipErrorInvalidValue;
    info->n = kDft[id].n;
    info->r1 = kDft[id].r1;
    info->r2 = kDft[id].r2;
    info->shape = id == 0 ? 32 : 16;
    info->tiled = kDft[id].tiled;
    return 0;
}

extern "C" long long ofx_mfma_kernel_const_bytes(int id, int mode) {
    const int P = parts_of(mode);
    if (P == 0 || id < 0 || id >= kDftCount) return 0;
    if (id == 0) return ofx_mfma_const_bytes(mode);
    const int S1 = kDft[id].r2 / 16, S2 = kDft[id].r1 / 16;
    return (long long)(S1 * S1 + S2 * S2) * P * 64 * 16;
}

extern "C" int ofx_gen_mfma_kernel_consts(int id, int mode, void* table, hipStream_t stream) {
    if (id == 0) return ofx_gen_mfma_consts(mode, table, stream);
    if (id < 0 || id >= kDftCount || parts_of(mode) == 0 || table == nullptr) return (int)hipErrorInvalidValue;
    const int S1 = kDft[id].r2 / 16, S2 = kDft[id].r1 / 16;
    const int total = (S1 * S1 + S2 * S2) * 64;
    const dim3 grid((unsigned)((total + 255) / 256)), block(256);
    switch (mode) {
        case OFX_MFMA_BF16:
            hipLaunchKernelGGL(k_gen_dft_consts<OFX_MFMA_BF16>, grid, block, 0, stream, table, kDft[id].r1, kDft[id].r2);
            break;
        case OFX_MFMA_FP16X2:
            hipLaunchKernelGGL(k_gen_dft_consts<OFX_MFMA_FP16X2>, grid, block, 0, stream, table, kDft[id].r1, kDft[id].r2);
            break;
        default:
            hipLaunchKernelGGL(k_gen_dft_consts<OFX_MFMA_BF16X3>, grid, block, 0, stream, table, kDft[id].r1, kDft[id].r2);
            break;
    }
    return (int)hipGetLastError();
}

extern "C" int ofx_launch_mfma_kernel(int id, int mode, const ofx_args* a, int blocks, hipStream_t stream) {
    if (a == nullptr || id < 0 || id >= kDftCount || parts_of(mode) == 0) return (int)hipErrorInvalidValue;
    if (id == 0)
        return (a->tiles != nullptr || (a->flags & OFX_FLAG_MFMA_2D) != 0) ? (int)hipErrorInvalidValue
                                                                           : ofx_launch_mfma_c2c(mode, 1024, a, blocks, stream);
    if (a->batch <= 0) return 0;
    const int n = kDft[id].n;
    const bool dense = kDft[id].r1 == 1;
    // dense kernels: 16 transforms per wave, no tile lists, no 2D form
    if (dense && (a->tiles != nullptr || (a->flags & OFX_FLAG_MFMA_2D) != 0)) return (int)hipErrorInvalidValue;
    const long long need = dense ? (a->batch + 63) / 64 : (a->batch + 3) / 4;
    // one wave per transform; by default a grid of 2048 workgroups strides over the batch (with a
    // tile list: one workgroup per tile, `batch` tiles)
    if (a->tiles != nullptr && a->batch > 0x7fffffffLL) return (int)hipErrorInvalidConfiguration;
    const unsigned grid = a->tiles != nullptr ? (unsigned)std::min<long long>(a->batch, blocks > 0 ? blocks : 1024)
                                              : (unsigned)std::min<long long>(need, blocks > 0 ? blocks : 2048);
    // streaming HBM access when input and output together exceed the 256 MB last-level cache by
    // half again (the threshold of the FP32 kernels' streaming variants; tile lists never stream)
    const bool nt = a->tiles == nullptr && a->batch * (long long)n * 16 >= (384LL << 20);
    const bool t2d = (a->flags & OFX_FLAG_MFMA_2D) != 0;
    if (t2d && a->tiles != nullptr) return (int)hipErrorInvalidValue;
    dft_kernel_t k = nullptr;
    switch (mode) {
        case OFX_MFMA_BF16: k = dft_kernel_for<OFX_MFMA_BF16>(id, nt, t2d); break;
        case OFX_MFMA_FP16X2: k = dft_kernel_for<OFX_MFMA_FP16X2>(id, nt, t2d); break;
        default: k = dft_kernel_for<OFX_MFMA_BF16X3>(id, nt, t2d); break;
    }
    if (k == nullptr) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kDftBlock), 0, stream, *a);
    return (int)hipGetLastError();
}
