// kernels_1d.hip - 1D FP32 Stockham kernels for gfx950 (C2C, R2C, C2R) and the kernel registry.
//
// Data layout in HBM: transform t of a batch starts at in + t*idist (complex64 units for complex
// data, float units for real data); elements are contiguous.  One workgroup owns TPB whole
// transforms, one launch covers the whole batch, HBM is read once and written once.
//
// R2C of even n = 2M (reference behaviour: src/rfft1_d/arch35/fft/rfft1_d_fft_kernel.h:39-228):
//   z[m] = x[2m] + i x[2m+1] is the real input reinterpreted as M complex values (no pack pass),
//   Z = FFT_M(z) stays in LDS, and the Hermitian unpack
//     X[k] = (Z[k] + conj Z[M-k])/2 - (i/2) W_n^k (Z[k] - conj Z[M-k]),   k = 0..M
//   is fused into the store (DC and Nyquist imaginary parts written as exact zeros).
// C2R of even n (reference: src/irfft1_d/arch35/fft/irfft1_d_fft_kernel.h:40-180 expands to n and
//   runs an n-point inverse; here the pre-pack below halves the work):
//     Z'[k] = (X[k] + conj X[M-k]) + i W_n^{-k} (X[k] - conj X[M-k]),      k = 0..M-1
//   with Im X[0] = Im X[M] = 0, followed by an M-point backward FFT whose complex output is the
//   interleaved real result (unnormalised, like the reference).
// Odd n: R2C runs the full n-point complex FFT of (x, 0) and keeps bins 0..n/2; C2R expands the
//   Hermitian half-spectrum to n points on load and keeps the real part.
#include <hip/hip_runtime.h>
#include "ofx_launch.h"
#include "stockham.h"

namespace ofx {

// ---------------------------------------------------------------------------------------------
// HBM functors (V = float2 for complex64 / float, double2 for complex128 / double)
// ---------------------------------------------------------------------------------------------
// The output scale of a launch in V's precision: complex128 kernels take ofx_args::scale64 when it
// is set (a float scale would round factors such as 1/3 to single precision).
template <class S>
__device__ __forceinline__ S arg_scale(const ofx_args& a) {
    if constexpr (sizeof(S) == 8) {
        return a.scale64 != 0.0 ? a.scale64 : (double)a.scale;
    } else {
        return a.scale;
    }
}

// NT: streaming (non-temporal) accesses.  Used only where every wave's accesses cover whole
// 128-byte lines (rows of a multiple of 128 bytes, contiguous): there they skip the cache
// retention of data touched once; on partial lines they turn into re-reads and masked writes.
template <bool NT, class P>
__device__ __forceinline__ auto ldx(const P* p) {
    if constexpr (NT) {
        return ld_nt(p);
    } else {
        return *p;
    }
}
template <bool NT, class P>
__device__ __forceinline__ void stx(P* p, P v) {
    if constexpr (NT) {
        st_nt(p, v);
    } else {
        *p = v;
    }
}
// rows of N elements of type E (the real or complex row of a kernel) cover whole 128-byte lines
template <class E, int N>
constexpr bool rows_whole_lines() {
    return (N * (int)sizeof(E)) % 128 == 0;
}

template <class V, bool NT = false>
struct GLoadC {  // complex input, conjugated for the backward direction
    const V* __restrict__ p;
    scalar_t<V> sg;
    __device__ __forceinline__ V operator()(int i) const {
        V v = ldx<NT>(p + i);
        return mkv(v.x, v.y * sg);
    }
    // elements i, i+1 (complex64, i even, 16-byte aligned rows)
    __device__ __forceinline__ float4 load2(int i) const {
        float4 v = ldx<NT>(reinterpret_cast<const float4*>(p + i));
        return make_float4(v.x, v.y * sg, v.z, v.w * sg);
    }
};
template <class V, bool NT = false>
struct GStoreC {
    V* __restrict__ p;
    scalar_t<V> sg, scale;
    bool ok;
    __device__ __forceinline__ void operator()(int i, V v) const {
        if (ok) stx<NT>(p + i, mkv(v.x * scale, v.y * (sg * scale)));
    }
    __device__ __forceinline__ void store2(int i, V a, V b) const {
        const scalar_t<V> si = sg * scale;
        if (ok) stx<NT>(reinterpret_cast<float4*>(p + i), make_float4(a.x * scale, a.y * si, b.x * scale, b.y * si));
    }
};

// C2R (even n) pre-pack, computed while loading pass 0.  Output is conjugated (backward via the
// forward kernel).
template <class V>
struct GLoadC2RPre {
    using S = scalar_t<V>;
    const V* __restrict__ X;   // M+1 bins of this transform
    const V* __restrict__ W;   // W_n^k, k = 0..M
    int M;
    S half_weight;             // 0.5 for the adjoint of R2C, 1 otherwise (bins 1..M-1)
    __device__ __forceinline__ V operator()(int k) const {
        V a = X[k];
        V b = X[M - k];
        if (k == 0) {
            a.y = S(0);
            b.y = S(0);
        } else {
            a = cscale(a, half_weight);
            b = cscale(b, half_weight);
        }
        V bc = cconj(b);
        V s = cadd(a, bc);
        V d = csub(a, bc);
        V w = W[k];
        // i * conj(w) * d
        V t = mul_posi(cmulc(d, w));
        V z = cadd(s, t);
        return cconj(z);
    }
    __device__ __forceinline__ float4 load2(int k) const {
        const V a = (*this)(k), b = (*this)(k + 1);
        return make_float4(a.x, a.y, b.x, b.y);
    }
};
// The same pre-pack with the bins staged in LDS first (X[0..M-1] at lds[pad(i)], one coalesced
// pass over HBM): the partner bin X[M-k] comes from LDS instead of a second, reversed global load.
// X[M] (the partner of bin 0 only) is read from HBM.
template <class C>
struct LdsC2RPre {
    using V = typename C::V;
    using S = scalar_t<V>;
    const V* lt;
    const V* __restrict__ X;
    const V* __restrict__ W;
    S half_weight;
    __device__ __forceinline__ V operator()(int k) const {
        constexpr int M = C::N;
        V a = lt[C::pad(k)];
        V b = (k == 0) ? X[M] : lt[C::pad(M - k)];
        if (k == 0) {
            a.y = S(0);
            b.y = S(0);
        } else {
            a = cscale(a, half_weight);
            b = cscale(b, half_weight);
        }
        const V bc = cconj(b);
        const V t = mul_posi(cmulc(csub(a, bc), W[k]));
        return cconj(cadd(cadd(a, bc), t));
    }
};

// odd n R2C: real input as complex with zero imaginary part; keep bins 0..n/2.
template <class V>
struct GLoadRealAsC {
    const scalar_t<V>* __restrict__ p;
    __device__ __forceinline__ V operator()(int i) const { return mkv(p[i], scalar_t<V>(0)); }
};
template <class V>
struct GStoreHalfSpectrum {
    using S = scalar_t<V>;
    V* __restrict__ p;
    int half;      // n/2
    S scale, weight;
    bool ok;
    __device__ __forceinline__ void operator()(int i, V v) const {
        if (ok && i <= half) {
            // odd n has no Nyquist bin: every bin but DC has a distinct conjugate partner
            const S s = (i == 0) ? scale : scale * weight;
            V r = mkv(v.x * s, v.y * s);
            if (i == 0) r.y = S(0);
            p[i] = r;
        }
    }
};
// odd n C2R: Hermitian expansion on load (conjugated for the backward direction), real part out.
template <class V>
struct GLoadHermExpand {
    const V* __restrict__ X;
    int n;
    scalar_t<V> half_weight;
    __device__ __forceinline__ V operator()(int i) const {
        const int h = n / 2;
        V v;
        if (i == 0) {
            v = mkv(X[0].x, scalar_t<V>(0));
        } else if (i <= h) {
            v = cscale(X[i], half_weight);
        } else {
            v = cconj(cscale(X[n - i], half_weight));
        }
        return cconj(v);
    }
};
template <class V>
struct GStoreReal {
    scalar_t<V>* __restrict__ p;
    scalar_t<V> scale;
    bool ok;
    __device__ __forceinline__ void operator()(int i, V v) const {
        if (ok) p[i] = v.x * scale;
    }
};

// ---------------------------------------------------------------------------------------------
// staged tile copies (HBM <-> LDS, 16-byte accesses): tile = `nvalid` contiguous transforms
// ---------------------------------------------------------------------------------------------
template <class C>
__device__ __forceinline__ void stage_in(typename C::V* lds, const typename C::V* __restrict__ src, int nvalid,
                                         scalar_t<typename C::V> sg) {
    using V = typename C::V;
    const int total = nvalid * C::N;
    if constexpr ((C::N & 1) == 0 && sizeof(V) == 8) {
        const float4* s4 = reinterpret_cast<const float4*>(src);
        for (int l = threadIdx.x; l < total / 2; l += C::BLOCK) {
            float4 v = ldx<C::NT_LOAD && rows_whole_lines<V, C::N>()>(s4 + l);
            const int e = 2 * l;
            const int t = e / C::N, i = e - t * C::N;
            V* d = lds + t * C::LDS_N;
            d[C::pad(i)] = mkv(v.x, v.y * sg);
            d[C::pad(i + 1)] = mkv(v.z, v.w * sg);
        }
    } else {
        for (int l = threadIdx.x; l < total; l += C::BLOCK) {
            V v = ldx<C::NT_LOAD && rows_whole_lines<V, C::N>()>(src + l);
            const int t = l / C::N, i = l - t * C::N;
            lds[t * C::LDS_N + C::pad(i)] = mkv(v.x, v.y * sg);
        }
    }
}
template <class C>
__device__ __forceinline__ void stage_out(const typename C::V* lds, typename C::V* __restrict__ dst, int nvalid,
                                          scalar_t<typename C::V> sg, scalar_t<typename C::V> scale) {
    using V = typename C::V;
    const int total = nvalid * C::N;
    const scalar_t<V> si = sg * scale;
    if constexpr ((C::N & 1) == 0 && sizeof(V) == 8) {
        float4* d4 = reinterpret_cast<float4*>(dst);
        for (int l = threadIdx.x; l < total / 2; l += C::BLOCK) {
            const int e = 2 * l;
            const int t = e / C::N, i = e - t * C::N;
            const V* s = lds + t * C::LDS_N;
            V a = s[C::pad(i)], b = s[C::pad(i + 1)];
            stx<C::NT_STORE && rows_whole_lines<V, C::N>()>(d4 + l, make_float4(a.x * scale, a.y * si, b.x * scale, b.y * si));
        }
    } else {
        for (int l = threadIdx.x; l < total; l += C::BLOCK) {
            const int t = l / C::N, i = l - t * C::N;
            V a = lds[t * C::LDS_N + C::pad(i)];
            stx<C::NT_STORE && rows_whole_lines<V, C::N>()>(dst + l, mkv(a.x * scale, a.y * si));
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Hermitian partners in one thread (IOV = 3: real transforms of even n = 2M)
// ---------------------------------------------------------------------------------------------
// The R2C unpack combines Z[i] with Z[M-i]; the C2R pre-pack combines X[i] with X[M-i].  In the
// last Stockham pass (Ns = M/R butterflies) butterfly j produces Z[j + k*Ns], k = 0..R-1, and the
// partner of j + k*Ns is (Ns - j) + (R-1-k)*Ns: output R-1-k of butterfly Ns - j.  The first pass
// (NBF = M/R butterflies, Ns = 1) reads its inputs at j + k*NBF, whose partners sit the same way
// in butterfly NBF - j.  So when one thread owns butterflies j and Ns - j (the "unit" j), the
// unpack / pre-pack runs in registers: the R2C stores its bins straight from the last pass, the C2R
// reads every bin from HBM once, in its first pass, and neither moves the spectrum through LDS an
// extra time.  Unit 0 owns butterfly 0, whose outputs pair among themselves (k with R-k), and for
// an even butterfly count butterfly Ns/2 (k with R-1-k).  (Ns + 1) / 2 units cover a pass.
template <bool NTS, class V>
__device__ __forceinline__ void r2c_pair_out(int i, int M, V zi, V zp, V w, V* __restrict__ out, scalar_t<V> ascale,
                                             scalar_t<V> wgt, bool ok) {
    // X[i] and X[M-i] from Z[i] and its partner Z[M-i] (Z[0] for i = 0, giving X[0] and X[M])
    using S = scalar_t<V>;
    const V zc = cconj(zp);
    const V e = cscale(cadd(zi, zc), S(0.5));
    const V o = cscale(csub(zi, zc), S(0.5));
    const V pk = cmul(w, o);
    const V lo = cadd(e, mul_negi(pk));
    const V hi = cconj(csub(e, mul_negi(pk)));
    if (!ok) return;
    if (i == 0) {
        stx<NTS>(out, mkv(lo.x * ascale, S(0)));
        stx<NTS>(out + M, mkv(hi.x * ascale, S(0)));
    } else {
        stx<NTS>(out + i, mkv(lo.x * wgt, lo.y * wgt));
        if (2 * i != M) stx<NTS>(out + (M - i), mkv(hi.x * wgt, hi.y * wgt));
    }
}

// C2R pre-pack value of bin k from X[k] (a) and X[M-k] (b), conjugated (backward via the forward
// kernel); the same formula as GLoadC2RPre.
template <class V>
__device__ __forceinline__ V c2r_prepack(int k, V a, V b, V w, scalar_t<V> half_weight) {
    using S = scalar_t<V>;
    if (k == 0) {
        a.y = S(0);
        b.y = S(0);
    } else {
        a = cscale(a, half_weight);
        b = cscale(b, half_weight);
    }
    const V bc = cconj(b);
    const V t = mul_posi(cmulc(csub(a, bc), w));
    return cconj(cadd(cadd(a, bc), t));
}

// passes 0 .. NP-2, the first reading `gin`, all writing this transform's LDS tile
template <class C, int P = 0, class GIn>
__device__ __forceinline__ void passes_into_lds(typename C::V* lt, int tid, const typename C::V* __restrict__ tw,
                                                const GIn& gin) {
    LdsRead<C> lr{lt};
    LdsWrite<C> lw{lt};
    if constexpr (P == 0)
        stockham_pass<C, 0, false, true>(tid, tw, gin, lw);
    else
        stockham_pass<C, P, true, true>(tid, tw, lr, lw);
    if constexpr (P + 2 < C::NP) passes_into_lds<C, P + 1>(lt, tid, tw, gin);
}

// R2C last pass with the unpack in registers: reads LDS, writes the M+1 bins to HBM.
template <class C>
__device__ __forceinline__ void r2c_last_pass_paired(const typename C::V* lt, int tid,
                                                     const typename C::V* __restrict__ tw,
                                                     const typename C::V* __restrict__ W, typename C::V* __restrict__ out,
                                                     scalar_t<typename C::V> ascale, scalar_t<typename C::V> wgt,
                                                     bool ok) {
    using V = typename C::V;
    constexpr int P = C::NP - 1;
    constexpr int R = C::radix(P);
    constexpr int Ns = C::ns(P);  // = M / R, the pass's butterfly count
    constexpr int M = C::N;
    constexpr int U = (Ns + 1) / 2;
    constexpr int NU = (U + C::T - 1) / C::T;
    static_assert(P > 0, "paired real kernels have at least two passes");
    const V* __restrict__ t0 = tw + C::tw_off(P);
#pragma unroll
    for (int r = 0; r < NU; ++r) {
        const int u = tid + r * C::T;
        if (U % C::T == 0 || u < U) {
            const bool hasb = (u != 0) || (Ns % 2 == 0);
            const int ja = u;
            const int jb = u ? Ns - u : Ns / 2;
            V va[R], vb[R];
#pragma unroll
            for (int k = 0; k < R; ++k) va[k] = lt[C::pad(ja + k * Ns)];
            if (hasb) {
#pragma unroll
                for (int k = 0; k < R; ++k) vb[k] = lt[C::pad(jb + k * Ns)];
            }
#pragma unroll
            for (int k = 1; k < R; ++k) va[k] = cmul(va[k], t0[ja * (R - 1) + k - 1]);
            Dft<R, V>::run(va);
            if (hasb) {
#pragma unroll
                for (int k = 1; k < R; ++k) vb[k] = cmul(vb[k], t0[jb * (R - 1) + k - 1]);
                Dft<R, V>::run(vb);
            }
            if (u != 0) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int i = ja + k * Ns;
                    r2c_pair_out<C::NT_STORE>(i, M, va[k], vb[R - 1 - k], W[i], out, ascale, wgt, ok);
                }
            } else {
#pragma unroll
                for (int k = 0; 2 * k <= R; ++k) r2c_pair_out<C::NT_STORE>(k * Ns, M, va[k], va[(R - k) % R], W[k * Ns], out, ascale, wgt, ok);
                if constexpr (Ns % 2 == 0) {
#pragma unroll
                    for (int k = 0; 2 * k <= R - 1; ++k) {
                        const int i = Ns / 2 + k * Ns;
                        r2c_pair_out<C::NT_STORE>(i, M, vb[k], vb[R - 1 - k], W[i], out, ascale, wgt, ok);
                    }
                }
            }
        }
    }
}

// C2R first pass with the pre-pack in registers: reads the M+1 bins from HBM, writes LDS.
template <class C>
__device__ __forceinline__ void c2r_first_pass_paired(typename C::V* lt, int tid, const typename C::V* __restrict__ X,
                                                      const typename C::V* __restrict__ W,
                                                      scalar_t<typename C::V> hw) {
    using V = typename C::V;
    constexpr int R = C::radix(0);
    constexpr int M = C::N;
    constexpr int NBF = M / R;
    constexpr int U = (NBF + 1) / 2;
    constexpr int NU = (U + C::T - 1) / C::T;
    static_assert(C::NP > 1, "paired real kernels have at least two passes");
#pragma unroll
    for (int r = 0; r < NU; ++r) {
        const int u = tid + r * C::T;
        if (U % C::T == 0 || u < U) {
            const bool hasb = (u != 0) || (NBF % 2 == 0);
            const int ja = u;
            const int jb = u ? NBF - u : NBF / 2;
            // the pre-pack replaces each partner pair in place: (xa[k], xb[R-1-k]) for u >= 1,
            // (xa[k], xa[R-k]) and (xb[k], xb[R-1-k]) for unit 0.  The partner's root is
            // W[M-i] = -conj(W[i]) (W = W_2M: W^M = -1), exactly so in the rounded table too, so
            // a pair loads one root.
            V xa[R], xb[R];
#pragma unroll
            for (int k = 0; k < R; ++k) xa[k] = X[ja + k * NBF];
            if (hasb) {
#pragma unroll
                for (int k = 0; k < R; ++k) xb[k] = X[jb + k * NBF];
            }
            if (u != 0) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int ia = ja + k * NBF, ib = jb + (R - 1 - k) * NBF;  // ib = M - ia
                    const V a = xa[k], b = xb[R - 1 - k];
                    const V w = W[ia];
                    xa[k] = c2r_prepack(ia, a, b, w, hw);
                    xb[R - 1 - k] = c2r_prepack(ib, b, a, mkv(-w.x, w.y), hw);
                }
            } else {
                const V a0 = xa[0];
                xa[0] = c2r_prepack(0, a0, X[M], W[0], hw);
#pragma unroll
                for (int k = 1; 2 * k <= R; ++k) {
                    const V a = xa[k], b = xa[R - k];
                    const V w = W[k * NBF];
                    xa[k] = c2r_prepack(k * NBF, a, b, w, hw);
                    if (2 * k != R) xa[R - k] = c2r_prepack((R - k) * NBF, b, a, mkv(-w.x, w.y), hw);
                }
                if constexpr (NBF % 2 == 0) {
#pragma unroll
                    for (int k = 0; 2 * k <= R - 1; ++k) {
                        const int ia = NBF / 2 + k * NBF, ib = NBF / 2 + (R - 1 - k) * NBF;
                        const V a = xb[k], b = xb[R - 1 - k];
                        const V w = W[ia];
                        xb[k] = c2r_prepack(ia, a, b, w, hw);
                        if (2 * k != R - 1) xb[R - 1 - k] = c2r_prepack(ib, b, a, mkv(-w.x, w.y), hw);
                    }
                }
            }
            Dft<R, V>::run(xa);
#pragma unroll
            for (int k = 0; k < R; ++k) lt[C::pad(ja * R + k)] = xa[k];
            if (hasb) {
                Dft<R, V>::run(xb);
#pragma unroll
                for (int k = 0; k < R; ++k) lt[C::pad(jb * R + k)] = xb[k];
            }
        }
    }
    C::sync();
}

// ---------------------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------------------
// SPLITX configurations (never staged) exchange between passes through a scalar LDS tile per
// transform, real parts then imaginary parts (split_fft): half the LDS of the complex tile.
template <class C>
__global__ __launch_bounds__(C::BLOCK) void k_c2c(ofx_args a) {
    using V = typename C::V;
    using S = scalar_t<V>;
    const S ascale = arg_scale<S>(a);
    static_assert(!(C::SPLITX && (C::STAGED || C::IOV != 1)), "split exchange is for direct single-element I/O");
    __shared__ std::conditional_t<C::SPLITX, S, V> lds[C::TPB * C::LDS_N];
    const int tl = threadIdx.x / C::T;
    const int tid = threadIdx.x - tl * C::T;
    const long long t0 = (long long)blockIdx.x * C::TPB;
    const long long t = t0 + tl;
    const V* __restrict__ tw = static_cast<const V*>(a.tw);
    if constexpr (C::SPLITX) {
        const bool ok = t < a.batch;
        const long long tt = ok ? t : 0;
        constexpr bool NTL = C::NT_LOAD && rows_whole_lines<V, C::N>();
        constexpr bool NTS = C::NT_STORE && rows_whole_lines<V, C::N>();
        GLoadC<V, NTL> gin{static_cast<const V*>(a.in) + tt * a.idist, a.in_sign};
        GStoreC<V, NTS> gout{static_cast<V*>(a.out) + tt * a.odist, a.out_sign, ascale, ok};
        if constexpr (C::WAVE_SYNC) {
            if (!ok) return;
        }
        split_fft<C>(lds + tl * C::LDS_N, tid, tw, gin, gout);
        return;
    } else {
    V* lt = lds + tl * C::LDS_N;
    if constexpr (C::STAGED) {
        // staged tiles require contiguous transforms (idist == odist == N): checked by the launcher
        const int nvalid = (int)min((long long)C::TPB, a.batch - t0);
        const V* src = static_cast<const V*>(a.in) + t0 * C::N;
        V* dst = static_cast<V*>(a.out) + t0 * C::N;
        stage_in<C>(lds, src, nvalid, a.in_sign);
        __syncthreads();
        LdsRead<C> gin{lt};
        LdsWrite<C> gout{lt};
        stockham_passes<C, true, true>(lt, tid, tw, gin, gout);
        __syncthreads();
        stage_out<C>(lds, dst, nvalid, a.out_sign, ascale);
    } else {
        const bool ok = t < a.batch;
        const long long tt = ok ? t : 0;
        constexpr bool NTL = C::NT_LOAD && rows_whole_lines<V, C::N>();
        constexpr bool NTS = C::NT_STORE && rows_whole_lines<V, C::N>();
        GLoadC<V, NTL> gin{static_cast<const V*>(a.in) + tt * a.idist, a.in_sign};
        GStoreC<V, NTS> gout{static_cast<V*>(a.out) + tt * a.odist, a.out_sign, ascale, ok};
        if constexpr (C::WAVE_SYNC) {
            if (!ok) return;  // whole transforms per wave: no block-wide barrier follows
        }
        stockham_passes<C, false, false>(lt, tid, tw, gin, gout);
    }
    }
}

// The weight of R2C bins 1..M-1: 2 for the adjoint of C2R (each stands for itself and its
// conjugate partner), 1 otherwise; the output scale included.
template <class S>
__device__ __forceinline__ S r2c_weight(const ofx_args& a, S ascale) {
    return ((a.flags & OFX_FLAG_HERMITIAN_WEIGHT) ? S(2) : S(1)) * ascale;
}

// Hermitian unpack of Z = FFT_M(z) held in LDS (lt[pad(k)]) into the M+1 bins at `out`, bins k and
// M-k together: with E = (Z[k] + conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / 2 and P = W^k O,
//     X[k] = E - iP,   X[M-k] = conj(E + iP)      (W^(M-k) = -conj(W^k), since W^M = -1)
// so each pair costs two LDS reads and one root of unity (k = 0 gives X[0] and X[M]).  DC and
// Nyquist imaginary parts are written as exact zeros.
template <class C>
__device__ __forceinline__ void r2c_unpack(const typename C::V* lt, int tid, const typename C::V* __restrict__ W,
                                           typename C::V* __restrict__ out, scalar_t<typename C::V> ascale,
                                           scalar_t<typename C::V> wgt) {
    using V = typename C::V;
    using S = scalar_t<V>;
    constexpr int M = C::N;
    for (int k = tid; k <= M / 2; k += C::T) {
        const V zk = lt[C::pad(k)];
        const V zc = cconj(lt[C::pad(k == 0 ? 0 : M - k)]);
        const V e = cscale(cadd(zk, zc), S(0.5));
        const V o = cscale(csub(zk, zc), S(0.5));
        const V pk = cmul(W[k], o);
        const V lo = cadd(e, mul_negi(pk));           // X[k]
        const V hi = cconj(csub(e, mul_negi(pk)));     // X[M-k] = conj(E + iP)
        if (k == 0) {
            stx<C::NT_STORE>(out, mkv(lo.x * ascale, S(0)));
            stx<C::NT_STORE>(out + M, mkv(hi.x * ascale, S(0)));
        } else {
            stx<C::NT_STORE>(out + k, mkv(lo.x * wgt, lo.y * wgt));
            if (2 * k != M) stx<C::NT_STORE>(out + (M - k), mkv(hi.x * wgt, hi.y * wgt));
        }
    }
}

// R2C, even n = 2*C::N.  in: real (idist floats), out: complex (odist complex, >= C::N+1).
template <class C>
__global__ __launch_bounds__(C::BLOCK, C::MINW) void k_r2c(ofx_args a) {
    using V = typename C::V;
    using S = scalar_t<V>;
    const S ascale = arg_scale<S>(a);
    __shared__ V lds[C::TPB * C::LDS_N];
    const int tl = threadIdx.x / C::T;
    const int tid = threadIdx.x - tl * C::T;
    const long long t0 = (long long)blockIdx.x * C::TPB;
    const long long t = t0 + tl;
    const bool ok = t < a.batch;
    const long long tt = ok ? t : 0;
    const V* __restrict__ tw = static_cast<const V*>(a.tw);
    V* lt = lds + tl * C::LDS_N;
    LdsWrite<C> lw{lt};
    if constexpr (C::IOV == 3) {
        if constexpr (C::WAVE_SYNC) {
            if (!ok) return;
        }
        // streaming variants: the real input rows (2N values) are read with streaming loads when
        // they are whole 128-byte lines; the bins are stored with streaming stores when NT_STORE
        constexpr bool NTI = C::NT_LOAD && rows_whole_lines<S, 2 * C::N>();
        GLoadC<V, NTI> gin{reinterpret_cast<const V*>(static_cast<const S*>(a.in) + tt * a.idist), 1.f};
        passes_into_lds<C>(lt, tid, tw, gin);
        r2c_last_pass_paired<C>(lt, tid, tw, static_cast<const V*>(a.tw_aux), static_cast<V*>(a.out) + tt * a.odist,
                                ascale, r2c_weight<S>(a, ascale), ok);
        return;
    } else if constexpr (C::STAGED) {
        const int nvalid = (int)min((long long)C::TPB, a.batch - t0);
        const V* src = reinterpret_cast<const V*>(static_cast<const S*>(a.in)) + t0 * C::N;
        stage_in<C>(lds, src, nvalid, S(1));
        __syncthreads();
        LdsRead<C> gin{lt};
        stockham_passes<C, true, true>(lt, tid, tw, gin, lw);
    } else {
        if constexpr (C::WAVE_SYNC) {
            if (!ok) return;
        }
        constexpr bool NTI = C::NT_LOAD && rows_whole_lines<S, 2 * C::N>();
        GLoadC<V, NTI> gin{reinterpret_cast<const V*>(static_cast<const S*>(a.in) + tt * a.idist), 1.f};
        stockham_passes<C, false, true>(lt, tid, tw, gin, lw);
    }
    if constexpr (!C::WAVE_SYNC || C::STAGED) __syncthreads();
    if (!ok) return;
    r2c_unpack<C>(lt, tid, static_cast<const V*>(a.tw_aux), static_cast<V*>(a.out) + tt * a.odist, ascale,
                  r2c_weight<S>(a, ascale));
}

// C2R, even n = 2*C::N.  in: complex (idist complex, >= C::N+1), out: real (odist floats).
template <class C>
__global__ __launch_bounds__(C::BLOCK, C::MINW) void k_c2r(ofx_args a) {
    using V = typename C::V;
    using S = scalar_t<V>;
    const S ascale = arg_scale<S>(a);
    __shared__ V lds[C::TPB * C::LDS_N];
    const int tl = threadIdx.x / C::T;
    const int tid = threadIdx.x - tl * C::T;
    const long long t0 = (long long)blockIdx.x * C::TPB;
    const long long t = t0 + tl;
    const bool ok = t < a.batch;
    const long long tt = ok ? t : 0;
    const V* __restrict__ tw = static_cast<const V*>(a.tw);
    V* lt = lds + tl * C::LDS_N;
    const S hw = (a.flags & OFX_FLAG_HERMITIAN_WEIGHT) ? S(0.5) : S(1);
    GLoadC2RPre<V> gin{static_cast<const V*>(a.in) + tt * a.idist, static_cast<const V*>(a.tw_aux), C::N, hw};
    if constexpr (C::IOV == 3) {
        if constexpr (C::WAVE_SYNC) {
            if (!ok) return;
        }
        // streaming-store variants: the real output rows are written with streaming stores
        constexpr bool NTO = C::NT_STORE;
        GStoreC<V, NTO> gout{reinterpret_cast<V*>(static_cast<S*>(a.out) + tt * a.odist), S(-1), ascale, ok};
        c2r_first_pass_paired<C>(lt, tid, static_cast<const V*>(a.in) + tt * a.idist, static_cast<const V*>(a.tw_aux),
                                 hw);
        LdsRead<C> lr{lt};
        stockham_passes<C, true, false, 1>(lt, tid, tw, lr, gout);
    } else if constexpr (C::STAGED) {
        const int nvalid = (int)min((long long)C::TPB, a.batch - t0);
        V* dst = reinterpret_cast<V*>(static_cast<S*>(a.out)) + t0 * C::N;
        LdsWrite<C> gout{lt};
        stockham_passes<C, false, true>(lt, tid, tw, gin, gout);
        __syncthreads();
        stage_out<C>(lds, dst, nvalid, S(-1), ascale);
    } else {
        if constexpr (C::WAVE_SYNC) {
            if (!ok) return;
        }
        // streaming-store variants: the real output rows are written with streaming stores
        constexpr bool NTO = C::NT_STORE;
        GStoreC<V, NTO> gout{reinterpret_cast<V*>(static_cast<S*>(a.out) + tt * a.odist), S(-1), ascale, ok};
        if constexpr (C::NP > 1) {
            // bins into LDS (coalesced), then the pre-pack reads both partners there
            const V* __restrict__ X = static_cast<const V*>(a.in) + tt * a.idist;
            constexpr int IT = (C::N + C::T - 1) / C::T;
            V xs[IT];  // every load in flight before the first LDS write
#pragma unroll
            for (int r = 0; r < IT; ++r)
                if (C::N % C::T == 0 || tid + r * C::T < C::N) xs[r] = X[tid + r * C::T];
#pragma unroll
            for (int r = 0; r < IT; ++r)
                if (C::N % C::T == 0 || tid + r * C::T < C::N) lt[C::pad(tid + r * C::T)] = xs[r];
            C::sync();
            LdsC2RPre<C> lin{lt, X, static_cast<const V*>(a.tw_aux), hw};
            stockham_passes<C, true, false>(lt, tid, tw, lin, gout);
        } else {
            stockham_passes<C, false, false>(lt, tid, tw, gin, gout);
        }
    }
}

// R2C, odd n = C::N.
template <class C>
__global__ __launch_bounds__(C::BLOCK) void k_r2c_odd(ofx_args a) {
    using V = typename C::V;
    using S = scalar_t<V>;
    const S ascale = arg_scale<S>(a);
    __shared__ V lds[C::TPB * C::LDS_N];
    const int tl = threadIdx.x / C::T;
    const int tid = threadIdx.x - tl * C::T;
    const long long t = (long long)blockIdx.x * C::TPB + tl;
    const bool ok = t < a.batch;
    const long long tt = ok ? t : 0;
    if constexpr (C::WAVE_SYNC) {
        if (!ok) return;
    }
    const S wgt = (a.flags & OFX_FLAG_HERMITIAN_WEIGHT) ? S(2) : S(1);
    GLoadRealAsC<V> gin{static_cast<const S*>(a.in) + tt * a.idist};
    GStoreHalfSpectrum<V> gout{static_cast<V*>(a.out) + tt * a.odist, C::N / 2, ascale, wgt, ok};
    stockham_passes<C, false, false>(lds + tl * C::LDS_N, tid, static_cast<const V*>(a.tw), gin, gout);
}

// C2R, odd n = C::N.
template <class C>
__global__ __launch_bounds__(C::BLOCK) void k_c2r_odd(ofx_args a) {
    using V = typename C::V;
    using S = scalar_t<V>;
    const S ascale = arg_scale<S>(a);
    __shared__ V lds[C::TPB * C::LDS_N];
    const int tl = threadIdx.x / C::T;
    const int tid = threadIdx.x - tl * C::T;
    const long long t = (long long)blockIdx.x * C::TPB + tl;
    const bool ok = t < a.batch;
    const long long tt = ok ? t : 0;
    if constexpr (C::WAVE_SYNC) {
        if (!ok) return;
    }
    const S hw = (a.flags & OFX_FLAG_HERMITIAN_WEIGHT) ? S(0.5) : S(1);
    GLoadHermExpand<V> gin{static_cast<const V*>(a.in) + tt * a.idist, C::N, hw};
    GStoreReal<V> gout{static_cast<S*>(a.out) + tt * a.odist, ascale, ok};
    stockham_passes<C, false, false>(lds + tl * C::LDS_N, tid, static_cast<const V*>(a.tw), gin, gout);
}

// Rows in, columns out: the row step of the four-step decomposition.  Transform t (length C::N,
// contiguous at in + t*idist) is written to out + img*odist + rhi*odist_hi + rlo + k*ostride with
// img = t / cols and t % cols = rhi*cols_lo + rlo (cols_lo = 0: rhi = 0).  The last pass stays in LDS and the workgroup stores its TPB transforms
// interleaved, so consecutive lanes write consecutive addresses (TPB*8-byte runs).  The LDS rows
// are RS = LDS_N + 1 elements apart (an odd count of 8-byte elements), so the TPB lanes reading
// element k of TPB different rows fall on distinct banks.
template <class C>
__global__ __launch_bounds__(C::BLOCK) void k_rowt(ofx_args a) {
    using V = typename C::V;
    using S = scalar_t<V>;
    const S ascale = arg_scale<S>(a);
    constexpr int RS = C::LDS_N + 1;
    constexpr int KSTEP = C::BLOCK / C::TPB;  // rows' elements stored per sweep of the workgroup
    static_assert(C::BLOCK % C::TPB == 0, "transposed store assumes TPB divides the block");
    __shared__ V lds[C::TPB * RS];
    const int tl = threadIdx.x / C::T;
    const int tid = threadIdx.x - tl * C::T;
    const long long t0 = (long long)blockIdx.x * C::TPB;
    const long long t = t0 + tl;
    const long long tt = t < a.batch ? t : 0;
    V* lt = lds + tl * RS;
    GLoadC<V> gin{static_cast<const V*>(a.in) + tt * a.idist, a.in_sign};
    LdsWrite<C> lw{lt};
    stockham_passes<C, false, true>(lt, tid, static_cast<const V*>(a.tw), gin, lw);
    __syncthreads();
    // thread -> (row r, first element k0); a row past the batch end stores nothing
    const int r = threadIdx.x % C::TPB;
    const int k0 = threadIdx.x / C::TPB;
    const long long tg = t0 + r;
    if (tg >= a.batch) return;
    const long long img = tg / a.cols;
    long long r_in = tg - img * a.cols, r_off = r_in;
    if (a.cols_lo > 0) {
        const long long rhi = r_in / a.cols_lo;
        r_off = rhi * a.odist_hi + (r_in - rhi * a.cols_lo);
    }
    V* __restrict__ o = static_cast<V*>(a.out) + img * a.odist + r_off + k0 * a.ostride;
    const long long step = (long long)KSTEP * a.ostride;
    const V* lr = lds + r * RS;
    const S si = a.out_sign * ascale;
#pragma unroll 4
    for (int k = k0; k < C::N; k += KSTEP, o += step) {
        const V v = lr[C::pad(k)];
        *o = mkv(v.x * ascale, v.y * si);
    }
}

// C2C over a tile list: workgroup b runs tiles[b], `count` (<= TPB) contiguous transforms starting
// at complex offset `offset` of both in and out (the groups of an uneven batch, one launch per
// length).  Offsets may be odd, so only single-element (8-byte) I/O configurations are built.
template <class C>
__global__ __launch_bounds__(C::BLOCK) void k_c2c_tiles(ofx_args a) {
    using V = typename C::V;
    using S = scalar_t<V>;
    const S ascale = arg_scale<S>(a);
    static_assert(C::IOV == 1 && !C::STAGED, "tile offsets carry no 16-byte alignment");
    __shared__ std::conditional_t<C::SPLITX, S, V> lds[C::TPB * C::LDS_N];
    const int tl = threadIdx.x / C::T;
    const int tid = threadIdx.x - tl * C::T;
    const ofx_batch_tile tile = static_cast<const ofx_batch_tile*>(a.tiles)[blockIdx.x];
    const bool ok = tl < tile.count;
    if constexpr (C::WAVE_SYNC) {
        if (!ok) return;  // whole transforms per wave: no block-wide barrier follows
    }
    const long long base = tile.offset + (long long)(ok ? tl : 0) * C::N;
    GLoadC<V> gin{static_cast<const V*>(a.in) + base, a.in_sign};
    GStoreC<V> gout{static_cast<V*>(a.out) + base, a.out_sign, ascale, ok};
    if constexpr (C::SPLITX)
        split_fft<C>(lds + tl * C::LDS_N, tid, static_cast<const V*>(a.tw), gin, gout);
    else
        stockham_passes<C, false, false>(lds + tl * C::LDS_N, tid, static_cast<const V*>(a.tw), gin, gout);
}

// ---------------------------------------------------------------------------------------------
// launchers + registry
// ---------------------------------------------------------------------------------------------
typedef hipError_t (*launch_fn)(const ofx_args*, hipStream_t);

// The kernel family of a kind: complex128 kinds run the same templates as their complex64
// counterparts, with C::V = double2.
constexpr int base_kind(int k) {
    return k == OFX_KIND_Z2Z ? OFX_KIND_C2C
           : k == OFX_KIND_D2Z ? OFX_KIND_R2C
           : k == OFX_KIND_Z2D ? OFX_KIND_C2R
           : k == OFX_KIND_D2Z_ODD ? OFX_KIND_R2C_ODD
           : k == OFX_KIND_Z2D_ODD ? OFX_KIND_C2R_ODD
           : k == OFX_KIND_ROWT_Z ? OFX_KIND_ROWT
                                   : k;
}
constexpr bool is_z_kind(int k) { return k >= OFX_KIND_Z2Z && k <= OFX_KIND_ROWT_Z; }

template <class C, int KIND_>
hipError_t launch(const ofx_args* a, hipStream_t s) {
    static_assert(is_z_kind(KIND_) == (sizeof(typename C::V) == 16), "kind and element type disagree");
    constexpr int KIND = base_kind(KIND_);
    if (a->batch <= 0) return hipSuccess;
    const long long blocks = KIND == OFX_KIND_C2C_TILES ? a->batch : (a->batch + C::TPB - 1) / C::TPB;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    if constexpr (C::IOV == 2 || (C::STAGED && (C::N % 2) == 0)) {
        // 16-byte accesses: every row must start 16-byte aligned
        if (KIND == OFX_KIND_C2C && ((a->idist | a->odist) & 1)) return hipErrorInvalidValue;
    }
    if constexpr (C::STAGED) {
        if (KIND == OFX_KIND_C2C && (a->idist != C::N || a->odist != C::N)) return hipErrorInvalidValue;
        if (KIND == OFX_KIND_R2C && a->idist != 2 * C::N) return hipErrorInvalidValue;
        if (KIND == OFX_KIND_C2R && a->odist != 2 * C::N) return hipErrorInvalidValue;
    }
    if constexpr (KIND == OFX_KIND_C2C) hipLaunchKernelGGL(k_c2c<C>, dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a);
    if constexpr (KIND == OFX_KIND_R2C) hipLaunchKernelGGL(k_r2c<C>, dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a);
    if constexpr (KIND == OFX_KIND_C2R) hipLaunchKernelGGL(k_c2r<C>, dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a);
    if constexpr (KIND == OFX_KIND_R2C_ODD)
        hipLaunchKernelGGL(k_r2c_odd<C>, dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a);
    if constexpr (KIND == OFX_KIND_C2R_ODD)
        hipLaunchKernelGGL(k_c2r_odd<C>, dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a);
    if constexpr (KIND == OFX_KIND_C2C_TILES) {
        if (a->tiles == nullptr) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_c2c_tiles<C>, dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a);
    }
    if constexpr (KIND == OFX_KIND_ROWT) {
        if (a->cols <= 0) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_rowt<C>, dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a);
    }
    return hipGetLastError();
}

struct Entry {
    ofx_kernel_info info;
    launch_fn fn;
};

template <class C, int KIND>
Entry make_entry() {
    Entry e{};
    e.info.kind = KIND;
    e.info.n = C::N;
    e.info.threads_per_transform = C::T;
    e.info.transforms_per_block = C::TPB;
    e.info.block_threads = C::BLOCK;
    e.info.npass = C::NP;
    for (int p = 0; p < C::NP; ++p) e.info.radices[p] = C::radix(p);
    e.info.staged = (C::STAGED || C::IOV == 2) ? 1 : 0;
    e.info.iov = C::IOV;
    e.info.lds_bytes = (int)(C::TPB * (C::LDS_N + (KIND == OFX_KIND_ROWT ? 1 : 0)) *
                             (C::SPLITX ? sizeof(scalar_t<typename C::V>) : sizeof(typename C::V)));
    static_assert(!C::SPLITX || KIND == OFX_KIND_C2C || KIND == OFX_KIND_C2C_TILES, "split exchange: C2C kernels only");
    e.info.tw_len = C::tw_len();
    e.info.stream = C::STREAM;
    e.fn = &launch<C, KIND>;
    return e;
}

#define OFX_CFG(N, T, TPB, PADSH, STG, ...) StockhamCfg<N, T, TPB, PADSH, STG, Radices<__VA_ARGS__>>
#define OFX_E(KIND, N, T, TPB, PADSH, STG, ...) make_entry<OFX_CFG(N, T, TPB, PADSH, STG, __VA_ARGS__), KIND>()
#define OFX_CFGV(N, T, TPB, PADSH, ...) StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, 2>
#define OFX_EV(KIND, N, T, TPB, PADSH, ...) make_entry<OFX_CFGV(N, T, TPB, PADSH, __VA_ARGS__), KIND>()
#define OFX_CFGZ(N, T, TPB, PADSH, ...) StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, 1, 0, 1, double2>
#define OFX_EZ(KIND, N, T, TPB, PADSH, ...) make_entry<OFX_CFGZ(N, T, TPB, PADSH, __VA_ARGS__), KIND>()
// split-exchange variant of a direct complex64 configuration
// real transforms with the Hermitian partners in one thread (IOV = 3), and their streaming variants
#define OFX_EH(KIND, N, T, TPB, PADSH, ...) make_entry<StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, 3>, KIND>()
// a configuration with I/O form IOV and streaming mode MODE (StockhamCfg::STREAM)
#define OFX_EM(KIND, N, T, TPB, PADSH, IOV, MODE, ...) \
    make_entry<StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, IOV, 0, 1, float2, MODE>, KIND>()
// the same for complex128
#define OFX_EZM(KIND, N, T, TPB, PADSH, MODE, ...) \
    make_entry<StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, 1, 0, 1, double2, MODE>, KIND>()
#define OFX_EX(KIND, N, T, TPB, PADSH, ...) \
    make_entry<StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, 1, 0, 1, float2, false, true>, KIND>()
// streaming variants (STREAM_ = true) of a direct, a paired and a complex128 configuration
#define OFX_ES(KIND, N, T, TPB, PADSH, ...) \
    make_entry<StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, 1, 0, 1, float2, true>, KIND>()
#define OFX_EVS(KIND, N, T, TPB, PADSH, ...) \
    make_entry<StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, 2, 0, 1, float2, true>, KIND>()
#define OFX_EZS(KIND, N, T, TPB, PADSH, ...) \
    make_entry<StockhamCfg<N, T, TPB, PADSH, false, Radices<__VA_ARGS__>, 1, 0, 1, double2, true>, KIND>()

#include "registry_1d.inc"

static const int kNumEntries = (int)(sizeof(kRegistry) / sizeof(kRegistry[0]));

// ---------------------------------------------------------------------------------------------
// twiddle tables: FP64 sincospi on the GPU, rounded once to FP32
// ---------------------------------------------------------------------------------------------
__global__ void k_gen_pass_tw(float2* tw, int off, int ns, int r) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int cnt = ns * (r - 1);
    if (idx >= cnt) return;
    const long long m = idx / (r - 1);
    const long long k = idx % (r - 1) + 1;
    const long long L = (long long)ns * r;
    const long long e = (m * k) % L;
    double s, c;
    sincospi(2.0 * (double)e / (double)L, &s, &c);
    tw[off + idx] = make_float2((float)c, (float)(-s));
}

__global__ void k_gen_roots(float2* tw, long long len, long long count) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const long long e = idx % len;
    double s, c;
    sincospi(2.0 * (double)e / (double)len, &s, &c);
    tw[idx] = make_float2((float)c, (float)(-s));
}

// complex128 tables: the same FP64 values, not rounded
__global__ void k_gen_pass_tw_z(double2* tw, int off, int ns, int r) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int cnt = ns * (r - 1);
    if (idx >= cnt) return;
    const long long m = idx / (r - 1);
    const long long k = idx % (r - 1) + 1;
    const long long L = (long long)ns * r;
    const long long e = (m * k) % L;
    double s, c;
    sincospi(2.0 * (double)e / (double)L, &s, &c);
    tw[off + idx] = make_double2(c, -s);
}

__global__ void k_gen_roots_z(double2* tw, long long len, long long count) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const long long e = idx % len;
    double s, c;
    sincospi(2.0 * (double)e / (double)len, &s, &c);
    tw[idx] = make_double2(c, -s);
}

}  // namespace ofx

using namespace ofx;

extern "C" int ofx_kernel_count(void) { return kNumEntries; }

extern "C" int ofx_kernel_info_get(int id, ofx_kernel_info* info) {
    if (id < 0 || id >= kNumEntries || info == nullptr) return (int)hipErrorInvalidValue;
    *info = kRegistry[id].info;
    return 0;
}

extern "C" int ofx_launch(int id, const ofx_args* args, hipStream_t stream) {
    if (id < 0 || id >= kNumEntries || args == nullptr) return (int)hipErrorInvalidValue;
    return (int)kRegistry[id].fn(args, stream);
}

extern "C" int ofx_gen_pass_twiddles(void* tw, int npass, const int* radices, hipStream_t stream) {
    if (npass < 0 || npass > OFX_MAX_PASSES) return (int)hipErrorInvalidValue;
    int ns = 1, off = 0;
    for (int p = 0; p < npass; ++p) {
        const int r = radices[p];
        if (p > 0) {
            const int cnt = ns * (r - 1);
            if (cnt > 0) {
                hipLaunchKernelGGL(k_gen_pass_tw, dim3((cnt + 255) / 256), dim3(256), 0, stream,
                                   static_cast<float2*>(tw), off, ns, r);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) return (int)e;
            }
            off += cnt;
        }
        ns *= r;
    }
    return 0;
}

extern "C" int ofx_gen_roots(void* tw, long long len, long long count, hipStream_t stream) {
    if (count <= 0) return 0;
    if (len <= 0) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gen_roots, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream,
                       static_cast<float2*>(tw), len, count);
    return (int)hipGetLastError();
}

extern "C" int ofx_gen_pass_twiddles_z(void* tw, int npass, const int* radices, hipStream_t stream) {
    if (npass < 0 || npass > OFX_MAX_PASSES) return (int)hipErrorInvalidValue;
    int ns = 1, off = 0;
    for (int p = 0; p < npass; ++p) {
        const int r = radices[p];
        if (p > 0) {
            const int cnt = ns * (r - 1);
            if (cnt > 0) {
                hipLaunchKernelGGL(k_gen_pass_tw_z, dim3((cnt + 255) / 256), dim3(256), 0, stream,
                                   static_cast<double2*>(tw), off, ns, r);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) return (int)e;
            }
            off += cnt;
        }
        ns *= r;
    }
    return 0;
}

extern "C" int ofx_gen_roots_z(void* tw, long long len, long long count, hipStream_t stream) {
    if (count <= 0) return 0;
    if (len <= 0) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gen_roots_z, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream,
                       static_cast<double2*>(tw), len, count);
    return (int)hipGetLastError();
}
