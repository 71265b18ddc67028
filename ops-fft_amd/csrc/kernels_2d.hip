// kernels_2d.hip - one-pass 2D C2C for images that fit in LDS (nx, ny in {32, 64, 128}: the
// reference's 2D domain, src/fft2_d/arch32/dd/fft2_d_dd.cpp:17-174, where it runs Y = P X Q as two
// matrix products in one kernel).
//
// One workgroup per image: the image is staged into LDS with 16-byte loads, the nx row transforms
// (length ny, T1 lanes each, rows inside a wave) run LDS to LDS, then the ny column transforms
// (length nx, T2 lanes each, consecutive lanes on consecutive columns) run from LDS and their last
// pass stores straight to HBM, one contiguous row segment per wave instruction.  HBM is touched
// once per element in each direction (16 B/point), where the row kernel + column kernel pair moves
// 32 B/point.
//
// LDS layout: element (r, j) at r*RS + j + (j >> 4): each row padded as the 1D kernels pad a
// transform, and the row stride RS = ny + ny/16 + 1 odd, so that the column passes (whose lanes
// address consecutive columns of one row) and the row passes (four to eight lanes per row, rows
// RS apart) spread over the banks.
#include <hip/hip_runtime.h>

#include <atomic>

#include "ofx_launch.h"
#include "stockham.h"

namespace ofx {

template <int NX_, int NY_, int T1_, class RowRadices, class ColRadices>
struct Fused2D {
    static constexpr int NX = NX_, NY = NY_;
    static constexpr int PADSH = 4;
    static constexpr int RS = NY + (NY >> PADSH) + 1;  // row stride in LDS (odd)
    static constexpr int BLOCK = NX * T1_;
    static constexpr int T2 = BLOCK / NY;
    // rows: NX transforms of length NY, T1 lanes each (TPB = NX sizes nothing here: the row bases
    // are placed by hand at r * RS)
    using RowC = StockhamCfg<NY, T1_, NX, PADSH, false, RowRadices>;
    // columns: element i of a column at i * RS from the column's padded position in a row
    using ColC = StockhamCfg<NX, T2, 1, PADSH, false, ColRadices, 1, NY, RS - NY>;
    // the same with LDS-only barriers, for the pipelined kernel (its next image's loads stay in
    // flight through the column passes)
    using ColCP = StockhamCfg<NX, T2, 1, PADSH, false, ColRadices, 1, NY, RS - NY, float2, 0, false, true>;
    // float4 (element pair) loads per lane for one image
    static constexpr int NV = NX * NY / 2 / BLOCK;
    // the pipelined form pays off where the one-image kernel's row passes (radix-16 rows, ny >= 64)
    // already hold it to 4-5 waves per SIMD; the 32-point rows run at 8 waves per SIMD one image
    // per workgroup, and the prefetch registers would halve that (2^26 points: 32^2 189 vs 173 us,
    // 64x32 185 vs 172 us, 128x32 188 vs 183 us pipelined)
    static constexpr bool PIPE_OK = (NX * NY / 2) % BLOCK == 0 && NY >= 64;
    // waves per SIMD the pipelined kernel's registers must allow: all the workgroups its LDS tile
    // lets a CU hold (160 KiB), at most 4 (128 VGPRs: the prefetch and the passes without spills)
    static constexpr int PIPE_WGS = 163840 / (NX * RS * 8) > 0 ? 163840 / (NX * RS * 8) : 1;
    static constexpr int PIPE_MINW = PIPE_WGS * (BLOCK / 64) / 4 > 4 ? 4 : PIPE_WGS * (BLOCK / 64) / 4;
    static_assert(T2 * NY == BLOCK && BLOCK <= 1024, "row and column phases share the workgroup");
    static_assert(RowC::WAVE_SYNC, "a row's lanes must stay inside one wave");
    static constexpr int LDS = NX * RS;
    __device__ __forceinline__ static int pos(int r, int j) { return r * RS + j + (j >> PADSH); }
};

// column store: element i of column c of the image
template <class C, bool NTS>
struct Store2D {
    using V = typename C::V;
    V* __restrict__ p;  // element (0, c) of the output image
    int stride;
    float sg, scale;
    __device__ __forceinline__ void operator()(int i, V v) const {
        if constexpr (NTS)
            st_nt(p + i * stride, mkv(v.x * scale, v.y * (sg * scale)));
        else
            p[i * stride] = mkv(v.x * scale, v.y * (sg * scale));
    }
};

// args: in/out images of NX*NY complex64 at idist/odist apart (batch images), tw = the row passes'
// twiddle table, tw_aux = the column passes' table
template <class F, bool NTL, bool NTS>
__global__ __launch_bounds__(F::BLOCK) void k_c2c_2d_lds(ofx_args a) {
    using RowC = typename F::RowC;
    using ColC = typename F::ColC;
    __shared__ float2 lds[F::LDS];
    const long long img = blockIdx.x;
    const float2* __restrict__ src = static_cast<const float2*>(a.in) + img * a.idist;
    float2* __restrict__ dst = static_cast<float2*>(a.out) + img * a.odist;
    // stage in: 16-byte loads of element pairs (j, j+1), j even, which never straddle a pad (j+1
    // is odd); images that do not start 16-byte aligned take 8-byte loads
    {
        const float sg = a.in_sign;
        if ((reinterpret_cast<uintptr_t>(src) & 15) == 0) {
            const float4* s4 = reinterpret_cast<const float4*>(src);
            for (int l = threadIdx.x; l < F::NX * F::NY / 2; l += F::BLOCK) {
                float4 v;
                if constexpr (NTL)
                    v = ld_nt(s4 + l);
                else
                    v = s4[l];
                const int e = 2 * l;
                const int r = e / F::NY, j = e - r * F::NY;
                float2* d = lds + F::pos(r, j);
                d[0] = make_float2(v.x, v.y * sg);
                d[1] = make_float2(v.z, v.w * sg);
            }
        } else {
            for (int e = threadIdx.x; e < F::NX * F::NY; e += F::BLOCK) {
                const float2 v = src[e];
                const int r = e / F::NY, j = e - r * F::NY;
                lds[F::pos(r, j)] = make_float2(v.x, v.y * sg);
            }
        }
    }
    __syncthreads();
    // rows, LDS to LDS
    {
        const int r = threadIdx.x / RowC::T;
        const int tid = threadIdx.x - r * RowC::T;
        float2* row = lds + r * F::RS;
        const LdsRead<RowC> rd{row};
        const LdsWrite<RowC> wr{row};
        stockham_passes<RowC, true, true>(row, tid, static_cast<const float2*>(a.tw), rd, wr);
    }
    __syncthreads();
    // columns, LDS to HBM
    {
        const int c = threadIdx.x % F::NY;
        const int tid = threadIdx.x / F::NY;
        float2* col = lds + c + (c >> F::PADSH);
        const LdsRead<ColC> rd{col};
        const Store2D<ColC, NTS> st{dst + c, F::NY, a.out_sign, a.scale};
        stockham_passes<ColC, true, false>(col, tid, static_cast<const float2*>(a.tw_aux), rd, st);
    }
}

// Pipelined form (an image's load, passes and stores would otherwise run one after the other in a
// workgroup, and these sizes hold few workgroups per CU): a grid of resident workgroups, each
// taking images b, b + G, ...; the next image's 16-byte loads are issued into
// registers right after the current image is in LDS and stay in flight through its row and column
// passes (whose barriers order LDS only), and its column stores drain while the next image is
// written to LDS.  Images must start 16-byte aligned (checked by the launcher).
template <class F, bool NTL, bool NTS>
__global__ __launch_bounds__(F::BLOCK, F::PIPE_MINW) void k_c2c_2d_lds_pipe(ofx_args a) {
    using RowC = typename F::RowC;
    using ColC = typename F::ColCP;
    static_assert(F::PIPE_OK, "whole element pairs per lane");
    __shared__ float2 lds[F::LDS];
    const long long G = gridDim.x;
    long long img = blockIdx.x;
    const float sg = a.in_sign;
    float4 buf[F::NV];
    // buffer loads: the image base in a scalar resource, one lane offset for all NV loads and the
    // load index in the scalar offset, so the prefetch costs its data registers and one address
    auto load = [&](long long m) {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float2*>(static_cast<const float2*>(a.in) + m * a.idist), (short)0, F::NX * F::NY * 8,
            0x00020000);
        const int vo = (int)threadIdx.x * 16;
#pragma unroll
        for (int v = 0; v < F::NV; ++v)
            buf[v] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, vo, v * F::BLOCK * 16, NTL ? 2 : 0));
    };
    if (img < a.batch) load(img);
    for (; img < a.batch; img += G) {
#pragma unroll
        for (int v = 0; v < F::NV; ++v) {
            const int e = 2 * (int)(threadIdx.x + v * F::BLOCK);
            const int r = e / F::NY, j = e - r * F::NY;
            float2* d = lds + F::pos(r, j);
            d[0] = make_float2(buf[v].x, buf[v].y * sg);
            d[1] = make_float2(buf[v].z, buf[v].w * sg);
        }
        lds_barrier();
        if (img + G < a.batch) load(img + G);  // in flight through this image's passes
        {
            const int r = threadIdx.x / RowC::T;
            const int tid = threadIdx.x - r * RowC::T;
            float2* row = lds + r * F::RS;
            const LdsRead<RowC> rd{row};
            const LdsWrite<RowC> wr{row};
            stockham_passes<RowC, true, true>(row, tid, static_cast<const float2*>(a.tw), rd, wr);
        }
        lds_barrier();
        {
            const int c = threadIdx.x % F::NY;
            const int tid = threadIdx.x / F::NY;
            float2* col = lds + c + (c >> F::PADSH);
            const LdsRead<ColC> rd{col};
            const Store2D<ColC, NTS> st{static_cast<float2*>(a.out) + img * a.odist + c, F::NY, a.out_sign, a.scale};
            stockham_passes<ColC, true, false>(col, tid, static_cast<const float2*>(a.tw_aux), rd, st);
        }
        lds_barrier();  // every column read of LDS precedes the next image's writes
    }
}

// resident workgroups of the pipelined kernel on the current device (cached per device)
template <class F, bool NTL, bool NTS>
static long long pipe_slots() {
    static std::atomic<int> cache[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        dev = -1;
    }
    int slots = dev >= 0 ? cache[dev].load(std::memory_order_relaxed) : 0;
    if (slots <= 0) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_c2c_2d_lds_pipe<F, NTL, NTS>, F::BLOCK, 0) !=
                hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev < 0 ? 0 : dev) != hipSuccess ||
            per_cu <= 0 || cus <= 0) {
            (void)hipGetLastError();
            per_cu = 1;
            cus = 256;
        }
        slots = per_cu * cus;
        if (dev >= 0) cache[dev].store(slots, std::memory_order_relaxed);
    }
    return slots;
}

// The pipelined form runs when every resident workgroup gets at least 2 images: 128^2 x 512 images
// 28.4 vs 32.6 us, x 1024 51.3 vs 62.8 us, x 4096 201-203 vs 259-260 us; 64x64 x 16384 188 vs 201
// us; 64x128 / 128x64 x 8192 187 / 186 vs 223 / 226 us; one image per workgroup about the same
// (128^2 x 256: 15.4 vs 15.0 us).
template <class F, bool NTL, bool NTS>
static bool launch_2d_pipe(const ofx_args* a, hipStream_t s) {
    const long long slots = pipe_slots<F, NTL, NTS>();
    if (a->batch < 2 * slots) return false;
    hipLaunchKernelGGL((k_c2c_2d_lds_pipe<F, NTL, NTS>), dim3((unsigned)slots), dim3(F::BLOCK), 0, s, *a);
    return true;
}

template <class F>
hipError_t launch_2d(const ofx_args* a, hipStream_t s) {
    if (a->batch <= 0) return hipSuccess;
    if (a->batch > 0x7fffffffLL) return hipErrorInvalidConfiguration;
    const bool ntl = (a->flags & OFX_FLAG_NT_LOAD) != 0, nts = (a->flags & OFX_FLAG_NT_STORE) != 0;
    if constexpr (F::PIPE_OK) {
        // the pipelined form, for 16-byte aligned images
        const bool aligned = ((reinterpret_cast<uintptr_t>(a->in) | reinterpret_cast<uintptr_t>(a->out)) & 15) == 0 &&
                             ((a->idist | a->odist) & 1) == 0;
        if (aligned && (a->flags & OFX_FLAG_NO_PIPELINE) == 0) {
            const bool ran = ntl && nts ? launch_2d_pipe<F, true, true>(a, s)
                             : nts      ? launch_2d_pipe<F, false, true>(a, s)
                             : ntl      ? launch_2d_pipe<F, true, false>(a, s)
                                        : launch_2d_pipe<F, false, false>(a, s);
            if (ran) return hipGetLastError();
        }
    }
    if (ntl && nts)
        hipLaunchKernelGGL((k_c2c_2d_lds<F, true, true>), dim3((unsigned)a->batch), dim3(F::BLOCK), 0, s, *a);
    else if (nts)
        hipLaunchKernelGGL((k_c2c_2d_lds<F, false, true>), dim3((unsigned)a->batch), dim3(F::BLOCK), 0, s, *a);
    else if (ntl)
        hipLaunchKernelGGL((k_c2c_2d_lds<F, true, false>), dim3((unsigned)a->batch), dim3(F::BLOCK), 0, s, *a);
    else
        hipLaunchKernelGGL((k_c2c_2d_lds<F, false, false>), dim3((unsigned)a->batch), dim3(F::BLOCK), 0, s, *a);
    return hipGetLastError();
}

struct Entry2D {
    ofx_2d_kernel_info info;
    hipError_t (*fn)(const ofx_args*, hipStream_t);
};

template <class F>
Entry2D entry_2d() {
    Entry2D e{};
    e.info.nx = F::NX;
    e.info.ny = F::NY;
    e.info.block_threads = F::BLOCK;
    e.info.lds_bytes = (int)(F::LDS * sizeof(float2));
    e.info.row_npass = F::RowC::NP;
    for (int p = 0; p < F::RowC::NP; ++p) e.info.row_radices[p] = F::RowC::radix(p);
    e.info.row_tw_len = F::RowC::tw_len();
    e.info.col_npass = F::ColC::NP;
    for (int p = 0; p < F::ColC::NP; ++p) e.info.col_radices[p] = F::ColC::radix(p);
    e.info.col_tw_len = F::ColC::tw_len();
    e.fn = &launch_2d<F>;
    return e;
}

// (nx, ny, lanes per row, row radices, column radices): 32 = 8 x 4, 64 = 16 x 4, 128 = 16 x 8;
// lanes per row 4 for ny <= 64 and 8 for ny = 128, so the first pass's butterflies fill them
#define OFX_2D(NX, NY, T1, RR, RC) entry_2d<Fused2D<NX, NY, T1, RR, RC>>()
typedef Radices<8, 4> R32;
typedef Radices<16, 4> R64;
typedef Radices<16, 8> R128;
static const Entry2D kReg2D[] = {
    OFX_2D(32, 32, 4, R32, R32),   OFX_2D(32, 64, 4, R64, R32),   OFX_2D(32, 128, 8, R128, R32),
    OFX_2D(64, 32, 4, R32, R64),   OFX_2D(64, 64, 4, R64, R64),   OFX_2D(64, 128, 8, R128, R64),
    OFX_2D(128, 32, 4, R32, R128), OFX_2D(128, 64, 4, R64, R128), OFX_2D(128, 128, 8, R128, R128),
};
#undef OFX_2D
static const int kNum2D = (int)(sizeof(kReg2D) / sizeof(kReg2D[0]));

}  // namespace ofx

using namespace ofx;

extern "C" int ofx_2d_kernel_count(void) { return kNum2D; }

extern "C" int ofx_2d_kernel_info_get(int id, ofx_2d_kernel_info* info) {
    if (info == nullptr || id < 0 || id >= kNum2D) return (int)hipErrorInvalidValue;
    *info = kReg2D[id].info;
    return 0;
}

extern "C" int ofx_launch_2d(int id, const ofx_args* args, hipStream_t stream) {
    if (args == nullptr || id < 0 || id >= kNum2D) return (int)hipErrorInvalidValue;
    return (int)kReg2D[id].fn(args, stream);
}
