// stockham.h - single-launch, LDS-resident Stockham autosort FFT for gfx950.
//
// One transform of length N is owned by T threads (T <= 64 keeps it inside one wave64, so passes
// synchronise with wave-level fences instead of s_barrier); a workgroup holds TPB transforms.
// Pass p of radix R with Ns = R_0*...*R_{p-1}:
//     for j in [0, N/R):  v_k = src[j + k*N/R] * W_{Ns*R}^{k*(j mod Ns)}   (k = 0..R-1)
//                         V   = DFT_R(v)
//                         dst[(j div Ns)*Ns*R + (j mod Ns) + k*Ns] = V_k
// which is the textbook Stockham autosort (the reference's mixed-radix stage,
// src/fft1_d/arch35/mix/fft1_d_c2c_mix_kernel.h:1108-1176, uses the same index map).  Pass 0 reads
// HBM directly (consecutive j -> consecutive addresses, so every load of a wave is one contiguous
// 512 B run) and the last pass writes HBM directly (Ns = N/R, so the store index is j + k*N/R, again
// contiguous); the passes in between exchange through LDS.  HBM is therefore touched exactly once
// per element in each direction.  Optionally ("staged") the whole workgroup tile is moved between
// HBM and LDS with 16-byte vector accesses first/last, which suits short transforms.
//
// LDS layout: transform t of the block at lds[t*LDS_N + pad(i)], pad(i) = i + (i >> PADSH).  The
// pad breaks the power-of-two strides of the early passes (index j*R + k) across the 64 banks.
//
// Twiddle table (per plan, HBM, L1/L2 resident): pass p >= 1 stores W_{Ns*R}^{m*k} at
// tw[tw_off(p) + m*(R-1) + (k-1)] for m in [0, Ns), k in [1, R).
#pragma once
#include <hip/hip_runtime.h>
#include "codelets.h"

namespace ofx {

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also waits for the wave's
// outstanding global loads and stores (its fence covers global memory); a kernel that keeps the
// next tile's loads or the last tile's stores in flight across a barrier uses this one instead.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int... Rs>
struct Radices {
    static constexpr int count = sizeof...(Rs);
    static constexpr int v[count] = {Rs...};
};

// IOV_ = 2: in the first and last pass a thread owns adjacent butterflies j, j+1, so its HBM loads
// and stores move two complex64 values (16 bytes) per instruction.
// IOV_ = 3 (R2C / C2R kernels only): in the last (R2C) or first (C2R) pass a thread owns the
// butterflies j and Ns - j that hold Hermitian partners, so the unpack / pre-pack runs in registers
// (kernels_1d.hip).
// COLS_ > 0: column kernels.  The workgroup holds COLS_ transforms (columns of a row-major matrix)
// interleaved in LDS, element i of column c at lds[c + i*(COLS_ + COLPAD_)], and consecutive
// lanes take consecutive columns, so each HBM row segment of COLS_ values is one contiguous access.
// V_: the element type, float2 (complex64) or double2 (complex128).
// STREAM_: the kernel's HBM input and/or output use streaming (non-temporal) accesses: 1 both,
// 2 the stores only, 3 the loads only (0 neither); chosen by the plan by working-set size, or
// where measured faster.
// SPLITX_: the LDS tile holds one scalar per element and each exchange between passes moves the
// real parts, then the imaginary parts (split_fft below).
// LDSBAR_: the passes' workgroup barriers order LDS only (lds_barrier), so that global loads a
// kernel keeps in flight across the passes (the next tile's data) are not waited for.
template <int N_, int T_, int TPB_, int PADSH_, bool STAGED_, class RL, int IOV_ = 1, int COLS_ = 0, int COLPAD_ = 1,
          class V_ = float2, int STREAM_ = 0, bool SPLITX_ = false, bool LDSBAR_ = false>
struct StockhamCfg {
    using V = V_;
    static constexpr int STREAM = STREAM_;                       // 0 none, 1 both, 2 stores, 3 loads
    static constexpr bool NT_LOAD = STREAM_ == 1 || STREAM_ == 3;   // streaming HBM loads
    static constexpr bool NT_STORE = STREAM_ == 1 || STREAM_ == 2;  // streaming HBM stores
    static constexpr bool SPLITX = SPLITX_;
    static constexpr int N = N_;
    static constexpr int T = T_;
    static constexpr int TPB = TPB_;
    static constexpr int PADSH = PADSH_;
    static constexpr bool STAGED = STAGED_;
    static constexpr int IOV = IOV_;
    static constexpr int NP = RL::count;
    static constexpr int COLS = COLS_;
    static constexpr int COLSTRIDE = COLS_ + COLPAD_;
    static constexpr int BLOCK = T * TPB;
    static constexpr bool WAVE_SYNC = (COLS_ == 0) && (T <= 64) && (64 % T == 0);

    static constexpr int radix(int p) { return RL::v[p]; }
    static constexpr int ns(int p) {
        int s = 1;
        for (int q = 0; q < p; ++q) s *= RL::v[q];
        return s;
    }
    static constexpr int tw_off(int p) {
        int o = 0;
        for (int q = 1; q < p; ++q) o += ns(q) * (RL::v[q] - 1);
        return o;
    }
    static constexpr int tw_len() { return tw_off(NP); }
    static constexpr int LDS_N = COLS_ > 0 ? N * COLSTRIDE : N + (N >> PADSH);
    static constexpr bool USES_LDS = STAGED || NP > 1;
    static constexpr int LDS_ELEMS = USES_LDS ? TPB * LDS_N : 1;
    // Waves per SIMD that the register allocation must leave room for (the real kernels' launch
    // bounds): for the Hermitian-pair kernels (IOV 3), as many as their LDS tiles let a CU hold
    // (160 KiB of LDS, 32 waves), so the pair pass's extra registers cannot cost occupancy; 1 (no
    // bound) otherwise.
    static constexpr int wgs_per_cu() {
        const int by_lds = 163840 / (LDS_ELEMS * (int)sizeof(V_));
        const int by_waves = 32 / ((BLOCK + 63) / 64);
        return by_lds < by_waves ? by_lds : by_waves;
    }
    static constexpr bool LDSBAR = LDSBAR_;
    static constexpr int MINW = IOV_ != 3 ? 1
                                : (wgs_per_cu() * ((BLOCK + 63) / 64) / 4 > 8 ? 8
                                   : wgs_per_cu() * ((BLOCK + 63) / 64) / 4 < 1 ? 1
                                                                             : wgs_per_cu() * ((BLOCK + 63) / 64) / 4);
    static constexpr int max_elems_per_thread() {
        int m = 0;
        for (int p = 0; p < NP; ++p) {
            int r = RL::v[p];
            int nb = (N / r + T - 1) / T;
            if (nb * r > m) m = nb * r;
        }
        return m;
    }
    __device__ __forceinline__ static int pad(int i) {
        if constexpr (COLS_ > 0)
            return i * COLSTRIDE;
        else
            return i + (i >> PADSH);
    }
    __device__ __forceinline__ static void sync() {
        if constexpr (WAVE_SYNC) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else if constexpr (LDSBAR_) {
            lds_barrier();
        } else {
            __syncthreads();
        }
    }
    static_assert(N == ns(NP), "radix product must equal N");
    static_assert(IOV_ == 1 || sizeof(V_) == 8, "paired 16-byte I/O is for complex64 only");
};


// Reads element i of the current transform from LDS.
template <class C>
struct LdsRead {
    const typename C::V* base;
    __device__ __forceinline__ typename C::V operator()(int i) const { return base[C::pad(i)]; }
};
template <class C>
struct LdsWrite {
    typename C::V* base;
    __device__ __forceinline__ void operator()(int i, typename C::V v) const { base[C::pad(i)] = v; }
};

// One Stockham pass.  IN_LDS && OUT_LDS means source and destination are the same LDS buffer, so
// all reads must complete before any write.
template <class C, int P, bool IN_LDS, bool OUT_LDS, class Ld, class St>
__device__ __forceinline__ void stockham_pass(int tid, const typename C::V* __restrict__ tw, const Ld& ld,
                                              const St& st) {
    using V = typename C::V;
    constexpr int R = C::radix(P);
    constexpr int Ns = C::ns(P);
    constexpr int NBF = C::N / R;
    constexpr int NB = (NBF + C::T - 1) / C::T;
    constexpr bool EXACT = (NBF % C::T) == 0;
    // paired mapping (j = 2*tid + (b&1) + (b>>1)*2T) for 16-byte HBM access
    constexpr bool VLOAD = C::IOV == 2 && P == 0 && !IN_LDS && EXACT && NB % 2 == 0;
    constexpr bool VSTORE = C::IOV == 2 && P == C::NP - 1 && !OUT_LDS && EXACT && NB % 2 == 0 && Ns == NBF;
    constexpr bool PAIRED = VLOAD || VSTORE;
    auto jof = [tid](int b) { return PAIRED ? (2 * tid + (b & 1) + (b >> 1) * 2 * C::T) : (tid + b * C::T); };
    V v[NB][R];
    V w[NB][R > 1 ? R - 1 : 1];
    if constexpr (VLOAD) {
#pragma unroll
        for (int b = 0; b < NB; b += 2) {
            const int j = jof(b);
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const float4 q = ld.load2(j + k * NBF);
                v[b][k] = make_float2(q.x, q.y);
                v[b + 1][k] = make_float2(q.z, q.w);
            }
        }
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = jof(b);
        if (EXACT || j < NBF) {
            if constexpr (!VLOAD) {
#pragma unroll
                for (int k = 0; k < R; ++k) v[b][k] = ld(j + k * NBF);
            }
            if constexpr (P > 0) {
                const V* t = tw + C::tw_off(P) + (j % Ns) * (R - 1);
#pragma unroll
                for (int k = 0; k < R - 1; ++k) w[b][k] = t[k];
            }
        }
    }
    if constexpr (IN_LDS && OUT_LDS) C::sync();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = jof(b);
        if (EXACT || j < NBF) {
            if constexpr (P > 0) {
#pragma unroll
                for (int k = 1; k < R; ++k) v[b][k] = cmul(v[b][k], w[b][k - 1]);
            }
            Dft<R, V>::run(v[b]);
            if constexpr (!VSTORE) {
                const int base = (j / Ns) * (Ns * R) + (j % Ns);
#pragma unroll
                for (int k = 0; k < R; ++k) st(base + k * Ns, v[b][k]);
            }
        }
    }
    if constexpr (VSTORE) {
        // last pass: Ns = N/R, so butterfly j stores to j + k*Ns and j, j+1 are adjacent
#pragma unroll
        for (int b = 0; b < NB; b += 2) {
            const int j = jof(b);
#pragma unroll
            for (int k = 0; k < R; ++k) st.store2(j + k * Ns, v[b][k], v[b + 1][k]);
        }
    }
    if constexpr (OUT_LDS) C::sync();
}

// All passes.  The first pass reads `gin`, the last pass writes `gout`; SRC_LDS / DST_LDS state
// whether those functors address this transform's LDS buffer (staged tiles, fused pre/post
// processing) rather than HBM, which decides where the passes must synchronise.
template <class C, bool SRC_LDS, bool DST_LDS, int P = 0, class GIn, class GOut>
__device__ __forceinline__ void stockham_passes(typename C::V* lds_t, int tid, const typename C::V* __restrict__ tw,
                                                const GIn& gin, const GOut& gout) {
    constexpr bool FIRST = (P == 0);
    constexpr bool LAST = (P == C::NP - 1);
    LdsRead<C> lr{lds_t};
    LdsWrite<C> lw{lds_t};
    if constexpr (FIRST && LAST) {
        stockham_pass<C, P, SRC_LDS, DST_LDS>(tid, tw, gin, gout);
    } else if constexpr (FIRST) {
        stockham_pass<C, P, SRC_LDS, true>(tid, tw, gin, lw);
    } else if constexpr (LAST) {
        stockham_pass<C, P, true, DST_LDS>(tid, tw, lr, gout);
    } else {
        stockham_pass<C, P, true, true>(tid, tw, lr, lw);
    }
    if constexpr (!LAST) stockham_passes<C, SRC_LDS, DST_LDS, P + 1>(lds_t, tid, tw, gin, gout);
}

// ---- split exchange -------------------------------------------------------------------------
// The same passes with the exchange between pass P and P+1 done in two halves through a scalar
// LDS tile: write the real parts of pass P's outputs, barrier, read the real parts of pass P+1's
// inputs, barrier, and the same for the imaginary parts.  Four barriers per exchange instead of
// two, and half the LDS: a 2048-point column group of 8 columns needs 72 KiB instead of 144 KiB,
// so two workgroups share a CU and one loads from HBM while the other computes.
template <class C, int P>
struct PassDims {
    static constexpr int R = C::radix(P);
    static constexpr int Ns = C::ns(P);
    static constexpr int NBF = C::N / R;
    static constexpr int NB = (NBF + C::T - 1) / C::T;
    static constexpr bool EXACT = (NBF % C::T) == 0;
};

// the registers of pass P: v[b][k] = input k of butterfly j = tid + b*T
template <class C, int P>
using PassRegs = typename C::V[PassDims<C, P>::NB][PassDims<C, P>::R];

template <class C, int P, class GOut>
__device__ __forceinline__ void split_passes(PassRegs<C, P>& v, int tid, scalar_t<typename C::V>* lds,
                                             const typename C::V* __restrict__ tw, const GOut& gout) {
    using V = typename C::V;
    using D = PassDims<C, P>;
#pragma unroll
    for (int b = 0; b < D::NB; ++b) {
        const int j = tid + b * C::T;
        if (D::EXACT || j < D::NBF) {
            if constexpr (P > 0) {
                const V* t = tw + C::tw_off(P) + (j % D::Ns) * (D::R - 1);
#pragma unroll
                for (int k = 1; k < D::R; ++k) v[b][k] = cmul(v[b][k], t[k - 1]);
            }
            Dft<D::R, V>::run(v[b]);
        }
    }
    if constexpr (P == C::NP - 1) {
#pragma unroll
        for (int b = 0; b < D::NB; ++b) {
            const int j = tid + b * C::T;
            if (D::EXACT || j < D::NBF) {
                const int base = (j / D::Ns) * (D::Ns * D::R) + (j % D::Ns);
#pragma unroll
                for (int k = 0; k < D::R; ++k) gout(base + k * D::Ns, v[b][k]);
            }
        }
    } else {
        using E = PassDims<C, P + 1>;
        PassRegs<C, P + 1> u;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
#pragma unroll
            for (int b = 0; b < D::NB; ++b) {
                const int j = tid + b * C::T;
                if (D::EXACT || j < D::NBF) {
                    const int base = (j / D::Ns) * (D::Ns * D::R) + (j % D::Ns);
#pragma unroll
                    for (int k = 0; k < D::R; ++k) lds[C::pad(base + k * D::Ns)] = half ? v[b][k].y : v[b][k].x;
                }
            }
            C::sync();
#pragma unroll
            for (int b = 0; b < E::NB; ++b) {
                const int j = tid + b * C::T;
                if (E::EXACT || j < E::NBF) {
#pragma unroll
                    for (int k = 0; k < E::R; ++k) {
                        const auto x = lds[C::pad(j + k * E::NBF)];
                        if (half)
                            u[b][k].y = x;
                        else
                            u[b][k].x = x;
                    }
                }
            }
            C::sync();
        }
        split_passes<C, P + 1>(u, tid, lds, tw, gout);
    }
}

// All passes of one column with the split exchange; `lds` is this column's scalar tile.
template <class C, class GIn, class GOut>
__device__ __forceinline__ void split_fft(scalar_t<typename C::V>* lds, int tid, const typename C::V* __restrict__ tw,
                                          const GIn& gin, const GOut& gout) {
    using D = PassDims<C, 0>;
    PassRegs<C, 0> v;
#pragma unroll
    for (int b = 0; b < D::NB; ++b) {
        const int j = tid + b * C::T;
        if (D::EXACT || j < D::NBF) {
#pragma unroll
            for (int k = 0; k < D::R; ++k) v[b][k] = gin(j + k * D::NBF);
        }
    }
    split_passes<C, 0>(v, tid, lds, tw, gout);
}

}  // namespace ofx
