// kernels_generic.hip - runtime-radix Stockham kernel for any length and layout, plus the
// elementwise kernels of large real transforms and the FP64 four-step twiddle tables.
//
// The templated kernels (kernels_1d.hip) cover the hot lengths with fully unrolled,
// register-resident passes.  This kernel covers everything else: the radix sequence is a kernel
// argument, each pass dispatches to the matching register codelet through a switch, and the data of
// the workgroup's transforms ping-pongs between two LDS buffers.  The same kernel serves strided
// layouts (columns of a 2D transform, the VERTICAL layout, both steps of the four-step
// decomposition) through the (t_hi, t_lo) addressing of ofx_generic_desc, with the load/store
// lane mapping chosen so that consecutive lanes touch consecutive addresses.
//
// Four-step twiddles W_N^e (N up to 2^27) are formed in FP64 from two FP64 tables,
// W_N^e = W_N^(e_hi*B) * W_N^e_lo with B = 2048, then rounded once to FP32 (kept in FP64 by the
// complex128 instantiation, which serves double-precision plans with FP64 pass tables and half
// the points per workgroup, since its LDS buffers hold 16-byte values).
#include <hip/hip_runtime.h>
#include <atomic>
#include "ofx_launch.h"
#include "codelets.h"
#include "twiddle64.h"

namespace ofx {

constexpr int kGenBlock = 256;
constexpr int kGenMaxTpb = 64;  // transforms per workgroup (ofx_generic_desc::tpb)
constexpr int kGenLoadU = 4;    // complex-input loads (and output elements) in flight per lane

struct GenParams {
    ofx_args a;
    ofx_generic_desc d;
};

// ---------------------------------------------------------------------------------------------
// one pass of radix R over all transforms of the workgroup (LDS src -> LDS dst)
// ---------------------------------------------------------------------------------------------
// LDS layout: transform t at buf + t*ln, element i at gpad(i) = i + i/16 (ln = gpad(n)).  The pad
// breaks the power-of-two strides of the passes (pass 0 writes element j*R + k from lane j: without
// it the lanes of a wave meet on a few banks).
__device__ __forceinline__ int gpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int gen_ln(int n) { return n + (n >> 4); }

template <int R, class V>
__device__ __forceinline__ void gen_pass(const V* __restrict__ src, V* __restrict__ dst, int n, int ln, int ns,
                                         int tpb, const V* __restrict__ tw) {
    const int nbf = n / R;
    const int total = tpb * nbf;
    for (int f = threadIdx.x; f < total; f += blockDim.x) {
        const int t = f / nbf;
        const int j = f - t * nbf;
        const V* s = src + t * ln;
        V* d = dst + t * ln;
        V v[R];
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = s[gpad(j + k * nbf)];
        const int m = j % ns;
        if (ns > 1) {
            const V* w = tw + m * (R - 1);
#pragma unroll
            for (int k = 1; k < R; ++k) v[k] = cmul(v[k], w[k - 1]);
        }
        Dft<R, V>::run(v);
        const int base = (j - m) * R + m;
#pragma unroll
        for (int k = 0; k < R; ++k) d[gpad(base + k * ns)] = v[k];
    }
}

// BIG = false: radices up to 16 only (the common smooth lengths; a much smaller register
// footprint, since the kernel's VGPR count is set by its largest codelet); BIG = true adds the
// prime radices 17..47.
template <bool BIG, class V>
__device__ __forceinline__ bool run_pass(int r, const V* src, V* dst, int n, int ns, int tpb, const V* tw) {
    const int ln = gen_ln(n);
    if constexpr (!BIG) {
        switch (r) {
            case 2: gen_pass<2, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 3: gen_pass<3, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 4: gen_pass<4, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 5: gen_pass<5, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 6: gen_pass<6, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 7: gen_pass<7, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 8: gen_pass<8, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 9: gen_pass<9, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 10: gen_pass<10, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 11: gen_pass<11, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 12: gen_pass<12, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 13: gen_pass<13, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 15: gen_pass<15, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 16: gen_pass<16, V>(src, dst, n, ln, ns, tpb, tw); return true;
            case 25:
                // complex64 only: the radix-25 codelet keeps the small form at 103 VGPRs (4 waves
                // per SIMD), where it would take complex128 from 131 to 171
                if constexpr (sizeof(V) == 8) {
                    gen_pass<25, V>(src, dst, n, ln, ns, tpb, tw);
                    return true;
                }
                return false;
            default: return false;
        }
    }
    switch (r) {
        case 2: gen_pass<2, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 3: gen_pass<3, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 4: gen_pass<4, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 5: gen_pass<5, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 6: gen_pass<6, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 7: gen_pass<7, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 8: gen_pass<8, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 9: gen_pass<9, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 10: gen_pass<10, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 11: gen_pass<11, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 12: gen_pass<12, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 13: gen_pass<13, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 15: gen_pass<15, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 16: gen_pass<16, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 17: gen_pass<17, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 19: gen_pass<19, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 23: gen_pass<23, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 25: gen_pass<25, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 29: gen_pass<29, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 31: gen_pass<31, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 37: gen_pass<37, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 41: gen_pass<41, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 43: gen_pass<43, V>(src, dst, n, ln, ns, tpb, tw); return true;
        case 47: gen_pass<47, V>(src, dst, n, ln, ns, tpb, tw); return true;
        default: return false;
    }
}


// ---------------------------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------------------------
template <bool BIG, class V>
__global__ __launch_bounds__(kGenBlock) void k_generic(GenParams p) {
    using S = scalar_t<V>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V* lds = reinterpret_cast<V*>(lds_raw);
    const ofx_args& a = p.a;
    const ofx_generic_desc& d = p.d;
    const int n = d.n;
    const int tpb = d.tpb;
    const int ln = gen_ln(n);
    V* buf0 = lds;
    V* buf1 = lds + (size_t)tpb * ln;
    const long long t0 = (long long)blockIdx.x * tpb;
    const int nvalid = (int)min((long long)tpb, a.batch - t0);
    const V* tw = static_cast<const V*>(a.tw);
    const S scale = (sizeof(S) == 8 && a.scale64 != 0.0) ? S(a.scale64) : S(a.scale);

    // per-transform addressing, once per workgroup: (t_hi, t_lo) needs a 64-bit division, which
    // per element would cost more than the element's arithmetic
    __shared__ long long s_ibase[kGenMaxTpb], s_obase[kGenMaxTpb], s_tlo[kGenMaxTpb];
    if (threadIdx.x < nvalid) {
        const long long tg = t0 + threadIdx.x;
        const long long thi = tg / d.cnt_lo, tlo = tg - thi * d.cnt_lo;
        s_ibase[threadIdx.x] = thi * a.idist + tlo * d.idist_lo;
        s_obase[threadIdx.x] = thi * a.odist + tlo * d.odist_lo;
        s_tlo[threadIdx.x] = tlo;
    }
    __syncthreads();

    // ---- load into buf0 ----------------------------------------------------------------------
    const int h = n / 2;  // used by the real-data modes
    // complex input (the common mode): kGenLoadU loads per thread issued before any is used, so
    // each lane keeps several HBM requests in flight (one at a time left the kernel bound by
    // memory latency at a quarter of the bandwidth)
    const bool cplx_in = d.in_mode == OFX_IN_COMPLEX;
    if (cplx_in) {
        const int total = nvalid * n;
        const int B = (int)blockDim.x;
        for (int f0 = threadIdx.x; f0 < total; f0 += B * kGenLoadU) {
            V vals[kGenLoadU];
            int dst_at[kGenLoadU];
#pragma unroll
            for (int u = 0; u < kGenLoadU; ++u) {
                const int f = f0 + u * B;
                dst_at[u] = -1;
                if (f < total) {
                    int t, i;
                    if (d.in_tmajor) {
                        t = f % nvalid;
                        i = f / nvalid;
                    } else {
                        t = f / n;
                        i = f - t * n;
                    }
                    vals[u] = static_cast<const V*>(a.in)[s_ibase[t] + i * a.istride];
                    dst_at[u] = t * ln + gpad(i);
                }
            }
#pragma unroll
            for (int u = 0; u < kGenLoadU; ++u) {
                if (dst_at[u] >= 0) {
                    V v = vals[u];
                    v.y *= S(a.in_sign);
                    buf0[dst_at[u]] = v;
                }
            }
        }
    }
    for (int f = cplx_in ? nvalid * n : threadIdx.x; f < nvalid * n; f += blockDim.x) {
        int t, i;
        if (d.in_tmajor) {
            t = f % nvalid;
            i = f / nvalid;
        } else {
            t = f / n;
            i = f - t * n;
        }
        const long long base = s_ibase[t];
        V v;
        if (d.in_mode == OFX_IN_COMPLEX) {
            v = static_cast<const V*>(a.in)[base + i * a.istride];
            v.y *= S(a.in_sign);
        } else if (d.in_mode == OFX_IN_REAL) {
            v = mkv(static_cast<const S*>(a.in)[base + i * a.istride], S(0));
        } else if (d.in_mode == OFX_IN_HERMITIAN) {
            // bins 0..h of an n-point Hermitian spectrum; conjugated for the backward transform
            const V* X = static_cast<const V*>(a.in) + base;
            const S hw = (a.flags & OFX_FLAG_HERMITIAN_WEIGHT) ? S(0.5) : S(1);
            if (i == 0) {
                v = mkv(X[0].x, S(0));
            } else if (i <= h) {
                v = cscale(X[(long long)i * a.istride], hw);
            } else {
                v = cconj(cscale(X[(long long)(n - i) * a.istride], hw));
            }
            v = cconj(v);
        } else {  // OFX_IN_C2R_PACK: n = M, spectra of 2M points with M+1 bins
            const V* X = static_cast<const V*>(a.in) + base;
            const V* W = static_cast<const V*>(a.tw_aux);
            const S hw = (a.flags & OFX_FLAG_HERMITIAN_WEIGHT) ? S(0.5) : S(1);
            V xa = X[(long long)i * a.istride];
            V xb = X[(long long)(n - i) * a.istride];
            if (i == 0) {
                xa.y = S(0);
                xb.y = S(0);
            } else {
                xa = cscale(xa, hw);
                xb = cscale(xb, hw);
            }
            const V bc = cconj(xb);
            const V s = cadd(xa, bc);
            const V df = csub(xa, bc);
            v = cconj(cadd(s, mul_posi(cmulc(df, W[i]))));
        }
        buf0[t * ln + gpad(i)] = v;
    }
    __syncthreads();

    // ---- passes -------------------------------------------------------------------------------
    V* src = buf0;
    V* dst = buf1;
    int ns = 1;
    for (int p = 0; p < d.npass; ++p) {
        const int r = d.radices[p];
        run_pass<BIG, V>(r, src, dst, n, ns, nvalid, tw + d.tw_off[p]);
        __syncthreads();
        ns *= r;
        V* tmp = src;
        src = dst;
        dst = tmp;
    }

    // ---- store from src ------------------------------------------------------------------------
    const double2* ftab = static_cast<const double2*>(a.tw_aux);
    const int nout = (d.out_mode == OFX_OUT_HALF) ? h + 1 : (d.out_mode == OFX_OUT_R2C_UNPACK ? n + 1 : n);
    // complex output (the common mode): kGenLoadU elements per lane, their LDS reads and four-step
    // twiddle loads issued before any is used
    const bool cplx_out = d.out_mode == OFX_OUT_COMPLEX;
    if (cplx_out) {
        const int total = nvalid * n;
        const int B = (int)blockDim.x;
        for (int f0 = threadIdx.x; f0 < total; f0 += B * kGenLoadU) {
            V vals[kGenLoadU], tws[kGenLoadU];
            long long at[kGenLoadU];
#pragma unroll
            for (int u = 0; u < kGenLoadU; ++u) {
                const int f = f0 + u * B;
                at[u] = -1;
                if (f < total) {
                    int t, k;
                    if (d.out_tmajor) {
                        t = f % nvalid;
                        k = f / nvalid;
                    } else {
                        t = f / n;
                        k = f - t * n;
                    }
                    vals[u] = src[t * ln + gpad(k)];
                    if (d.ft_len > 0) {
                        long long e = (s_tlo[t] + d.ft_base) * (long long)k;
                        if (e >= d.ft_len) e %= d.ft_len;  // never for the four-step's own steps
                        tws[u] = ft_twiddle_v((V*)nullptr, ftab, e, d.ft_len);
                    }
                    at[u] = s_obase[t] + k * a.ostride;
                }
            }
#pragma unroll
            for (int u = 0; u < kGenLoadU; ++u) {
                if (at[u] >= 0) {
                    V v = vals[u];
                    v.y *= S(a.out_sign);
                    if (d.ft_len > 0) v = cmul(v, tws[u]);
                    static_cast<V*>(a.out)[at[u]] = mkv(v.x * scale, v.y * scale);
                }
            }
        }
    }
    for (int f = cplx_out ? nvalid * nout : threadIdx.x; f < nvalid * nout; f += blockDim.x) {
        int t, k;
        if (d.out_tmajor) {
            t = f % nvalid;
            k = f / nvalid;
        } else {
            t = f / nout;
            k = f - t * nout;
        }
        const long long base = s_obase[t];
        const V* s = src + t * ln;
        if (d.out_mode == OFX_OUT_COMPLEX) {
            V v = s[gpad(k)];
            v.y *= S(a.out_sign);
            if (d.ft_len > 0) {
                long long e = (s_tlo[t] + d.ft_base) * (long long)k;
                if (e >= d.ft_len) e %= d.ft_len;  // never for the four-step's own steps
                v = cmul(v, ft_twiddle_v((V*)nullptr, ftab, e, d.ft_len));
            }
            static_cast<V*>(a.out)[base + k * a.ostride] = mkv(v.x * scale, v.y * scale);
        } else if (d.out_mode == OFX_OUT_HALF) {
            const S w = (k == 0) ? scale : scale * ((a.flags & OFX_FLAG_HERMITIAN_WEIGHT) ? S(2) : S(1));
            V v = cscale(s[gpad(k)], w);
            if (k == 0) v.y = S(0);
            static_cast<V*>(a.out)[base + k * a.ostride] = v;
        } else if (d.out_mode == OFX_OUT_REAL) {
            static_cast<S*>(a.out)[base + k * a.ostride] = s[gpad(k)].x * scale;
        } else {  // OFX_OUT_R2C_UNPACK: k = 0..n (n = M)
            const V* W = reinterpret_cast<const V*>(a.tw_aux);
            const int k1 = (k == n) ? 0 : k;
            const int k2 = (k == 0) ? 0 : n - k;
            const V zk = s[gpad(k1)];
            const V zc = cconj(s[gpad(k2)]);
            const V e = cscale(cadd(zk, zc), S(0.5));
            const V o = cscale(csub(zk, zc), S(0.5));
            V x = cadd(e, mul_negi(cmul(W[k], o)));
            S sc = scale;
            if (k == 0 || k == n) {
                x.y = S(0);
            } else if (a.flags & OFX_FLAG_HERMITIAN_WEIGHT) {
                sc *= S(2);
            }
            static_cast<V*>(a.out)[base + k * a.ostride] = mkv(x.x * sc, x.y * sc);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// FP64 four-step table: t[e] = W_len^e for e < B, t[B + h] = W_len^(h*B) for h < ceil(len/B)
// ---------------------------------------------------------------------------------------------
__global__ void k_gen_ft_table(double2* t, long long len, long long count) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    const long long e = (idx < kFtLo) ? idx : (idx - kFtLo) * kFtLo;
    double s, c;
    sincospi(2.0 * (double)(e % len) / (double)len, &s, &c);
    t[idx] = make_double2(c, -s);
}

// ---------------------------------------------------------------------------------------------
// large even-n real transforms: the M-point complex FFT runs as a four-step plan in a workspace
// and these kernels do the Hermitian unpack / pre-pack as separate passes.
// ---------------------------------------------------------------------------------------------
// Z: [batch][M] complex (workspace) -> X: [batch][M+1] complex
template <class V>
__global__ void k_r2c_unpack(const V* __restrict__ Z, V* __restrict__ X, const V* __restrict__ W, long long M,
                             long long batch, scalar_t<V> scale, int flags) {
    using S = scalar_t<V>;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = M + 1;
    if (idx >= per * batch) return;
    const long long t = idx / per, k = idx - t * per;
    const V* z = Z + t * M;
    const long long k1 = (k == M) ? 0 : k;
    const long long k2 = (k == 0) ? 0 : M - k;
    const V zk = z[k1];
    const V zc = cconj(z[k2]);
    const V e = cscale(cadd(zk, zc), S(0.5));
    const V o = cscale(csub(zk, zc), S(0.5));
    V x = cadd(e, mul_negi(cmul(W[k], o)));
    S sc = scale;
    if (k == 0 || k == M) {
        x.y = S(0);
    } else if (flags & OFX_FLAG_HERMITIAN_WEIGHT) {
        sc *= S(2);
    }
    X[t * per + k] = mkv(x.x * sc, x.y * sc);
}

// X: [batch][M+1] complex -> Z': [batch][M] complex, conjugated (the complex FFT that follows runs
// forward and conjugates its output)
template <class V>
__global__ void k_c2r_pack(const V* __restrict__ X, V* __restrict__ Z, const V* __restrict__ W, long long M,
                           long long batch, int flags) {
    using S = scalar_t<V>;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * batch) return;
    const long long t = idx / M, k = idx - t * M;
    const V* x = X + t * (M + 1);
    const S hw = (flags & OFX_FLAG_HERMITIAN_WEIGHT) ? S(0.5) : S(1);
    V xa = x[k], xb = x[M - k];
    if (k == 0) {
        xa.y = S(0);
        xb.y = S(0);
    } else {
        xa = cscale(xa, hw);
        xb = cscale(xb, hw);
    }
    const V bc = cconj(xb);
    Z[idx] = cconj(cadd(cadd(xa, bc), mul_posi(cmulc(csub(xa, bc), W[k]))));
}

}  // namespace ofx

using namespace ofx;

template <class V>
static int launch_r2c_unpack(const void* Z, void* X, const void* W, long long M, long long batch, scalar_t<V> scale,
                             int flags, hipStream_t stream) {
    const long long total = (M + 1) * batch;
    if (total <= 0) return 0;
    const long long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffLL) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_r2c_unpack<V>, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<const V*>(Z),
                       static_cast<V*>(X), static_cast<const V*>(W), M, batch, scale, flags);
    return (int)hipGetLastError();
}

template <class V>
static int launch_c2r_pack(const void* X, void* Z, const void* W, long long M, long long batch, int flags,
                           hipStream_t stream) {
    const long long total = M * batch;
    if (total <= 0) return 0;
    const long long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffLL) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_c2r_pack<V>, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<const V*>(X),
                       static_cast<V*>(Z), static_cast<const V*>(W), M, batch, flags);
    return (int)hipGetLastError();
}

extern "C" int ofx_r2c_unpack(const void* Z, void* X, const void* W, long long M, long long batch, float scale,
                              int flags, hipStream_t stream) {
    return launch_r2c_unpack<float2>(Z, X, W, M, batch, scale, flags, stream);
}

extern "C" int ofx_c2r_pack(const void* X, void* Z, const void* W, long long M, long long batch, int flags,
                            hipStream_t stream) {
    return launch_c2r_pack<float2>(X, Z, W, M, batch, flags, stream);
}

extern "C" int ofx_r2c_unpack_z(const void* Z, void* X, const void* W, long long M, long long batch, double scale,
                                int flags, hipStream_t stream) {
    return launch_r2c_unpack<double2>(Z, X, W, M, batch, scale, flags, stream);
}

extern "C" int ofx_c2r_pack_z(const void* X, void* Z, const void* W, long long M, long long batch, int flags,
                              hipStream_t stream) {
    return launch_c2r_pack<double2>(X, Z, W, M, batch, flags, stream);
}

extern "C" int ofx_generic_radix_supported(int r) {
    switch (r) {
        case 2: case 3: case 4: case 5: case 6: case 7: case 8: case 9: case 10: case 11: case 12: case 13:
        case 15: case 16: case 17: case 19: case 23: case 25: case 29: case 31: case 37: case 41: case 43:
        case 47:
            return 1;
        default:
            return 0;
    }
}

extern "C" int ofx_launch_generic(const ofx_args* args, const ofx_generic_desc* desc, hipStream_t stream) {
    if (args == nullptr || desc == nullptr) return (int)hipErrorInvalidValue;
    if (args->batch <= 0) return 0;
    const ofx_generic_desc& d = *desc;
    if (d.n < 1 || d.tpb < 1 || d.tpb > kGenMaxTpb || (long long)d.tpb * d.n > OFX_GENERIC_MAX_ELEMS || d.npass < 0 ||
        d.npass > OFX_MAX_PASSES || d.cnt_lo < 1)
        return (int)hipErrorInvalidValue;
    long long prod = 1;
    for (int p = 0; p < d.npass; ++p) {
        if (!ofx_generic_radix_supported(d.radices[p])) return (int)hipErrorInvalidValue;
        prod *= d.radices[p];
    }
    if (prod != d.n) return (int)hipErrorInvalidValue;
    if (d.ft_len > 0 && d.out_mode != OFX_OUT_COMPLEX) return (int)hipErrorInvalidValue;
    const long long blocks = (args->batch + d.tpb - 1) / d.tpb;
    if (blocks > 0x7fffffffLL) return (int)hipErrorInvalidConfiguration;
    const bool z = d.z != 0;
    if (z && (long long)d.tpb * d.n > OFX_GENERIC_MAX_ELEMS_Z) return (int)hipErrorInvalidValue;
    const size_t esz = z ? sizeof(double2) : sizeof(float2);
    const size_t lds = 2 * esz * (size_t)d.tpb * gen_ln(d.n);
    // the dynamic-LDS ceiling is a per-device function attribute: set once on each device
    static std::atomic<unsigned long long> attr_set{0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return (int)hipErrorInvalidDevice;
    if ((attr_set.load() & (1ULL << dev)) == 0) {
        // both element types share one LDS ceiling: 9216 complex64 or 4608 complex128 points
        const int max_lds = (int)(2 * sizeof(float2) * gen_ln(OFX_GENERIC_MAX_ELEMS));
        const void* fns[4] = {reinterpret_cast<const void*>(&k_generic<false, float2>),
                              reinterpret_cast<const void*>(&k_generic<true, float2>),
                              reinterpret_cast<const void*>(&k_generic<false, double2>),
                              reinterpret_cast<const void*>(&k_generic<true, double2>)};
        for (const void* f : fns) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
            if (e != hipSuccess) return (int)e;
        }
        attr_set.fetch_or(1ULL << dev);
    }
    bool big = false;
    for (int q = 0; q < d.npass; ++q) big = big || (d.radices[q] > 16 && (z || d.radices[q] != 25));
    GenParams p{*args, d};
    if (z) {
        if (big)
            hipLaunchKernelGGL((k_generic<true, double2>), dim3((unsigned)blocks), dim3(kGenBlock), lds, stream, p);
        else
            hipLaunchKernelGGL((k_generic<false, double2>), dim3((unsigned)blocks), dim3(kGenBlock), lds, stream, p);
    } else {
        if (big)
            hipLaunchKernelGGL((k_generic<true, float2>), dim3((unsigned)blocks), dim3(kGenBlock), lds, stream, p);
        else
            hipLaunchKernelGGL((k_generic<false, float2>), dim3((unsigned)blocks), dim3(kGenBlock), lds, stream, p);
    }
    return (int)hipGetLastError();
}

extern "C" int ofx_gen_ft_table(void* t, long long len, hipStream_t stream) {
    if (len <= 0) return (int)hipErrorInvalidValue;
    const long long count = kFtLo + (len + kFtLo - 1) / kFtLo;
    hipLaunchKernelGGL(k_gen_ft_table, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream,
                       static_cast<double2*>(t), len, count);
    return (int)hipGetLastError();
}

extern "C" long long ofx_ft_table_count(long long len) { return kFtLo + (len + kFtLo - 1) / kFtLo; }
