// kernels_col.hip - column transforms of row-major data (2D second pass, the VERTICAL layout,
// four-step column steps), complex64 and complex128.
//
// Data: `images` matrices of rows x cols complex values, row-major; one transform per column, length
// rows, element i of column c of image m at in[m*idist + c + i*istride] (istride = cols).  A
// workgroup takes COLS adjacent columns of one image: lane l works on column l % COLS, so in every
// load or store the COLS lanes of a group touch one contiguous COLS*8-byte run of a row.  The
// Stockham passes are the same register codelets as the row kernels, exchanging through LDS with
// the columns interleaved (stockham.h, COLS > 0).  Column groups are assigned XCD-aware: the 8 XCDs
// each take a contiguous range of groups (blocks are dispatched round-robin over XCDs), so the
// neighbouring groups that share HBM lines run on the same XCD at about the same time and meet in
// its L2.
#include <hip/hip_runtime.h>
#include "ofx_launch.h"
#include "stockham.h"
#include "twiddle64.h"
#include "twiddle_pow2.h"

namespace ofx {

template <class C>
struct ColLoad {
    using V = typename C::V;
    const V* __restrict__ p;  // element 0 of this column
    long long stride;
    scalar_t<V> sg;
    __device__ __forceinline__ V operator()(int i) const {
        V v = p[(long long)i * stride];
        return mkv(v.x, v.y * sg);
    }
};
// Factored register twiddles: when the last pass is exact (each of the T lanes of a column owns
// butterflies j = lane + b*T, and N/R is a multiple of T), every output index a lane stores is
// i = lane + T*m, m in [0, N/T), so W^(col*i) = W^(col*lane) * W^(col*T*m): one twiddle per lane
// and a per-column table of N/T entries in LDS (filled by the column's lanes before the passes)
// replace one twiddle evaluation per element.  Each factor carries sincospi's rounding, the product
// about 2 ulp.
template <class C>
struct FtFactored {
    static constexpr int E = C::N / C::T;
    // per-column stride of the LDS table: one spare entry, so that the COLS columns' entries m
    // fall in different banks (lanes of one column read the same m and share the access)
    static constexpr int ES = E + 1;
    static constexpr bool OK = (C::N % C::T == 0) && ((C::N / C::radix(C::NP - 1)) % C::T == 0);
};

// FT: 0 = none; 1 = four-step twiddle W_ftlen^(col * i) on output i from the FP64 two-level table;
// 2 = the same twiddle computed in registers (ftlen = 2^ftlog2); 3 = the factored register form
// (FtFactored) in the element type's precision; 4 = the factored form in FP64 for complex64 data
// (FP64 factors, FP64 product, one rounding: OFX_FLAG_FT_FP64).
template <class C, int FT>
struct ColStore {
    using V = typename C::V;
    V* __restrict__ p;
    long long stride;
    scalar_t<V> sg, scale;
    bool ok;
    const double2* __restrict__ ft;
    long long col, ftlen;
    int ftlog2;
    const V* ftab;  // FT 3: this column's W^(col*T*m), m < N/T
    V fbase;        // FT 3: W^(col*lane)
    int lane;
    const double2* ftab64;  // FT 4: the same in FP64
    double2 fbase64;
    __device__ __forceinline__ void operator()(int i, V v) const {
        if constexpr (FT == 1) v = cmul(v, ft_twiddle_v((V*)nullptr, ft, col * (long long)i, ftlen));
        if constexpr (FT == 2) v = cmul(v, twiddle_pow2_v((V*)nullptr, (unsigned long long)col * (unsigned)i, ftlog2));
        if constexpr (FT == 3) v = cmul(v, cmul(fbase, ftab[(i - lane) / C::T]));
        if constexpr (FT == 4) {
            const double2 w = cmul(fbase64, ftab64[(i - lane) / C::T]);
            v = cmul(v, mkv((scalar_t<V>)w.x, (scalar_t<V>)w.y));
        }
        if (ok) p[(long long)i * stride] = mkv(v.x * scale, v.y * (sg * scale));
    }
};

template <class C, int FT>
__global__ __launch_bounds__(C::BLOCK) void k_col(ofx_args a, long long cols, long long groups_per_image,
                                                  long long ftlen, int ftlog2) {
    using V = typename C::V;
    using S = scalar_t<V>;
    __shared__ std::conditional_t<C::SPLITX, S, V> lds[C::LDS_N];
    const int c = threadIdx.x % C::COLS;
    const int tid = threadIdx.x / C::COLS;
    // XCD-aware group index (blocks are dispatched round-robin over the 8 XCDs)
    const long long nblk = gridDim.x;
    long long g = blockIdx.x;
    if ((nblk & 7) == 0) g = (g & 7) * (nblk >> 3) + (g >> 3);
    const long long img = g / groups_per_image;
    const long long col = (g - img * groups_per_image) * C::COLS + c;
    const bool ok = col < cols;
    const long long cc = ok ? col : 0;
    const long long ccf = a.ft_div > 1 ? cc / a.ft_div : cc;  // the four-step twiddle's column index
    const S scale = (sizeof(S) == 8 && a.scale64 != 0.0) ? (S)a.scale64 : (S)a.scale;
    ColLoad<C> gin{static_cast<const V*>(a.in) + img * a.idist + cc, a.istride, a.in_sign};
    __shared__ V ftw[FT == 3 ? C::COLS * FtFactored<C>::ES : 1];
    __shared__ double2 ftw64[FT == 4 ? C::COLS * FtFactored<C>::ES : 1];
    V fbase = mkv(S(1), S(0));
    double2 fbase64 = make_double2(1.0, 0.0);
    if constexpr (FT == 4) {
        constexpr int E = FtFactored<C>::E;
        for (int m = tid; m < E; m += C::T)
            ftw64[c * FtFactored<C>::ES + m] = twiddle_pow2_z((unsigned long long)ccf * (unsigned)(C::T * m), ftlog2);
        fbase64 = twiddle_pow2_z((unsigned long long)ccf * (unsigned)tid, ftlog2);
        if constexpr (C::NP == 1) __syncthreads();  // otherwise the passes' barriers order the table
    }
    if constexpr (FT == 3) {
        constexpr int E = FtFactored<C>::E;
        for (int m = tid; m < E; m += C::T)
            ftw[c * FtFactored<C>::ES + m] = twiddle_pow2_v((V*)nullptr, (unsigned long long)ccf * (unsigned)(C::T * m), ftlog2);
        fbase = twiddle_pow2_v((V*)nullptr, (unsigned long long)ccf * (unsigned)tid, ftlog2);
        if constexpr (C::NP == 1) __syncthreads();  // otherwise the passes' barriers order the table
    }
    // output image offset; cols_lo > 0 splits the image index into (img / cols_lo) * odist_hi +
    // (img % cols_lo) * odist (the second step of a column four-step over several images)
    const long long oimg = a.cols_lo > 0 ? (img / a.cols_lo) * a.odist_hi + (img % a.cols_lo) * a.odist : img * a.odist;
    ColStore<C, FT> gout{static_cast<V*>(a.out) + oimg + cc, a.ostride, a.out_sign, scale, ok,
                         static_cast<const double2*>(a.tw_aux), ccf, ftlen, ftlog2,
                         ftw + (FT == 3 ? c * FtFactored<C>::ES : 0), fbase, tid,
                         ftw64 + (FT == 4 ? c * FtFactored<C>::ES : 0), fbase64};
    if constexpr (C::SPLITX)
        split_fft<C>(lds + c, tid, static_cast<const V*>(a.tw), gin, gout);
    else
        stockham_passes<C, false, false>(lds + c, tid, static_cast<const V*>(a.tw), gin, gout);
}

template <class C>
hipError_t launch_col(const ofx_args* a, hipStream_t s) {
    if (a->batch <= 0 || a->cols <= 0) return hipSuccess;
    const long long images = a->batch / a->cols;
    const long long gpi = (a->cols + C::COLS - 1) / C::COLS;
    const long long blocks = images * gpi;
    if (blocks > 0x7fffffffLL || images * a->cols != a->batch) return hipErrorInvalidValue;
    if (a->ft_len > 0) {
        const long long f = a->ft_len;
        int lg = 0;
        while ((1LL << lg) < f) ++lg;
        const bool fp64 = (a->flags & OFX_FLAG_FT_FP64) != 0 && sizeof(typename C::V) == 8;
        if ((1LL << lg) == f && lg >= 3 && lg <= 27 && !(fp64 && !FtFactored<C>::OK)) {
            constexpr int FT = FtFactored<C>::OK ? 3 : 2;
            constexpr int FT64 = (FtFactored<C>::OK && sizeof(typename C::V) == 8) ? 4 : FT;
            if (fp64)
                hipLaunchKernelGGL((k_col<C, FT64>), dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a, a->cols, gpi, f, lg);
            else
                hipLaunchKernelGGL((k_col<C, FT>), dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a, a->cols, gpi, f, lg);
        } else {
            // FP64 two-level table (any length; power-of-two lengths without the factored form
            // when FP64 twiddles are asked for)
            if (a->tw_aux == nullptr) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_col<C, 1>), dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a, a->cols, gpi, f, 0);
        }
    } else {
        hipLaunchKernelGGL((k_col<C, 0>), dim3((unsigned)blocks), dim3(C::BLOCK), 0, s, *a, a->cols, gpi, 0LL, 0);
    }
    return hipGetLastError();
}

struct ColEntry {
    ofx_kernel_info info;
    hipError_t (*fn)(const ofx_args*, hipStream_t);
};

template <class C>
ColEntry col_entry() {
    ColEntry e{};
    e.info.kind = sizeof(typename C::V) == 16 ? OFX_KIND_COL_Z : OFX_KIND_COL;
    e.info.n = C::N;
    e.info.threads_per_transform = C::T;
    e.info.transforms_per_block = C::COLS;
    e.info.block_threads = C::BLOCK;
    e.info.npass = C::NP;
    for (int p = 0; p < C::NP; ++p) e.info.radices[p] = C::radix(p);
    e.info.lds_bytes = (int)(C::LDS_N * (C::SPLITX ? sizeof(scalar_t<typename C::V>) : sizeof(typename C::V)));
    e.info.tw_len = C::tw_len();
    e.info.iov = 1;
    e.fn = &launch_col<C>;
    return e;
}

// (N, T, COLS, COLPAD, radices): T * COLS threads per workgroup
#define OFX_COL(N, T, COLS, COLPAD, ...) \
    col_entry<StockhamCfg<N, T, COLS, 4, false, Radices<__VA_ARGS__>, 1, COLS, COLPAD>>()
#define OFX_COLZ(N, T, COLS, COLPAD, ...) \
    col_entry<StockhamCfg<N, T, COLS, 4, false, Radices<__VA_ARGS__>, 1, COLS, COLPAD, double2>>()
// split (real / imaginary) LDS exchange: half the LDS per workgroup
#define OFX_COLX(N, T, COLS, COLPAD, ...) \
    col_entry<StockhamCfg<N, T, COLS, 4, false, Radices<__VA_ARGS__>, 1, COLS, COLPAD, float2, false, true>>()
#define OFX_COLZX(N, T, COLS, COLPAD, ...) \
    col_entry<StockhamCfg<N, T, COLS, 4, false, Radices<__VA_ARGS__>, 1, COLS, COLPAD, double2, false, true>>()

static const ColEntry kCols[] = {
    OFX_COL(2, 1, 64, 1, 2),
    OFX_COL(4, 1, 64, 1, 4),
    OFX_COL(8, 1, 64, 1, 8),
    OFX_COL(16, 1, 64, 1, 16),
    OFX_COL(32, 4, 32, 1, 4, 8),
    OFX_COL(64, 8, 16, 1, 8, 8),
    OFX_COL(64, 8, 32, 1, 8, 8),
    OFX_COL(128, 8, 16, 1, 8, 16),
    OFX_COL(128, 8, 32, 1, 8, 16),
    OFX_COL(256, 16, 16, 1, 16, 16),
    OFX_COL(256, 16, 32, 1, 16, 16),
    OFX_COL(512, 32, 16, 1, 8, 8, 8),
    OFX_COL(512, 32, 8, 1, 8, 8, 8),
    OFX_COL(1024, 64, 16, 1, 16, 16, 4),
    OFX_COL(1024, 64, 8, 1, 16, 16, 4),
    OFX_COL(1024, 128, 8, 1, 8, 8, 16),
    // split exchange, two workgroups per CU (58% vs 55% of HBM for 2048 x 8192 VERTICAL); radix
    // 8 x 16 x 16 first (2D 2048 x 2048 x 4: 93.1-94.2 us vs 95.6 us for 16 x 16 x 8); no padding
    // between the interleaved rows of 8 columns: never slower than one spare word, faster in some
    // processes (2D 2048 x 2048 x 4 89.1 vs 92.3 us in a process holding only its own buffers, a
    // tie at 91.0-91.6 us behind 1 GiB of other allocations; 2048 x 8192 VERTICAL 52.8-53.0 vs
    // 53.5-54.0 us; 2048 x 2048 x 8 a tie)
    OFX_COLX(2048, 64, 8, 0, 8, 16, 16),
    OFX_COLX(2048, 64, 8, 1, 32, 8, 8),
    OFX_COLX(2048, 64, 8, 1, 16, 16, 8),
    OFX_COL(2048, 128, 8, 1, 16, 16, 8),
    OFX_COL(2048, 128, 4, 1, 16, 16, 8),
    OFX_COL(2048, 256, 4, 1, 8, 8, 8, 4),
    OFX_COL(4096, 256, 4, 0, 16, 16, 16),
    OFX_COL(4096, 256, 2, 1, 16, 16, 16),
    // radix 4 x 16 x 16 first (64 x 1024^2 2D 392.7 vs 395.0 us, 2^20 x 64 four-step 483 vs 491 us,
    // 1024 x 65536 VERTICAL 275 vs 283 us for 16 x 8 x 8)
    OFX_COLX(1024, 32, 16, 1, 4, 16, 16),
    OFX_COLX(1024, 32, 16, 1, 16, 8, 8),
    // complex128 (16-byte elements: half the columns per workgroup for the same LDS)
    OFX_COLZ(2, 1, 64, 1, 2),
    OFX_COLZ(4, 1, 64, 1, 4),
    OFX_COLZ(8, 1, 64, 1, 8),
    OFX_COLZ(16, 1, 64, 1, 16),
    OFX_COLZ(32, 4, 32, 1, 4, 8),
    OFX_COLZ(64, 8, 16, 1, 8, 8),
    OFX_COLZ(128, 16, 16, 1, 8, 16),
    OFX_COLZ(256, 32, 16, 1, 4, 8, 8),
    OFX_COLZ(512, 64, 8, 1, 8, 8, 8),
    OFX_COLZ(1024, 128, 8, 1, 8, 8, 16),
    OFX_COLZ(2048, 256, 4, 0, 8, 8, 8, 4),
    OFX_COLZ(4096, 512, 2, 0, 8, 8, 8, 8),
    OFX_COLZX(1024, 64, 8, 1, 16, 16, 4),
    OFX_COLZX(2048, 128, 4, 1, 16, 16, 8),
};
static const int kNumCols = (int)(sizeof(kCols) / sizeof(kCols[0]));

}  // namespace ofx

using namespace ofx;

extern "C" int ofx_col_kernel_count(void) { return kNumCols; }

extern "C" int ofx_col_kernel_info_get(int id, ofx_kernel_info* info) {
    if (id < 0 || id >= kNumCols || info == nullptr) return (int)hipErrorInvalidValue;
    *info = kCols[id].info;
    return 0;
}

extern "C" int ofx_launch_col(int id, const ofx_args* args, hipStream_t stream) {
    if (id < 0 || id >= kNumCols || args == nullptr) return (int)hipErrorInvalidValue;
    return (int)kCols[id].fn(args, stream);
}
