// kernels_batched.hip - one persistent launch for a plan of many C2C transforms of uneven lengths.
//
// The plan (aclfftExtPlanManyUneven) holds groups g of b_g transforms of length n_g, stored one
// group after the other.  At plan time the groups are cut into tiles (a tile = the transforms one
// workgroup of the length's class processes together) and the tiles are sorted by decreasing cost
// (longest-processing-time order).  The kernel is launched with as many workgroups as fit on the
// device at once (ofx_batch_blocks_per_cu per CU); every workgroup repeatedly takes its next tile
// (from a device counter, or from a static round-robin over the sorted list), runs the Stockham
// body of that tile's length class, and exits when the queue is drained.  Each length class is a
// compiled StockhamCfg with 256 threads per workgroup, so one launch serves them all; the LDS
// buffer is sized for the largest class.  The counter is reset by a one-thread kernel
// launched just before (no memset node, so the pair can be captured in a hipGraph and replayed).
#include <hip/hip_runtime.h>

#include <atomic>

#include "ofx_launch.h"
#include "stockham.h"

namespace ofx {

constexpr int kBatchBlock = 256;

// Length classes, all with T * TPB = 256.  Two compiled class sets (the tuner times both):
// set 0 (the default) and set 1, which differs in the radix order / threads per transform of the
// 1000-, 1024-, 2048-, 4096- and 6561-point classes.
struct ClassSet0 {
    typedef StockhamCfg<60, 4, 64, 4, false, Radices<4, 3, 5>> C60;
    typedef StockhamCfg<128, 8, 32, 4, false, Radices<8, 16>> C128;
    typedef StockhamCfg<243, 16, 16, 30, false, Radices<27, 9>> C243;  // odd: no LDS pad needed
    typedef StockhamCfg<256, 16, 16, 4, false, Radices<16, 16>> C256;
    typedef StockhamCfg<512, 32, 8, 4, false, Radices<8, 8, 8>> C512;
    typedef StockhamCfg<1000, 128, 2, 4, false, Radices<10, 10, 10>> C1000;
    typedef StockhamCfg<1024, 64, 4, 4, false, Radices<8, 8, 16>> C1024;
    typedef StockhamCfg<2048, 128, 2, 4, false, Radices<16, 16, 8>> C2048;
    typedef StockhamCfg<4096, 256, 1, 4, false, Radices<16, 16, 16>> C4096;
    // 6561: split exchange (26 KiB of LDS instead of 52), so the kernel's LDS budget is set by the
    // 35 KiB classes and four workgroups fit on a CU
    typedef StockhamCfg<6561, 256, 1, 30, false, Radices<27, 27, 9>, 1, 0, 1, float2, false, true> C6561;
};
struct ClassSet1 {
    typedef ClassSet0::C60 C60;
    typedef ClassSet0::C128 C128;
    typedef ClassSet0::C243 C243;
    typedef ClassSet0::C256 C256;
    typedef ClassSet0::C512 C512;
    typedef StockhamCfg<1000, 128, 2, 4, false, Radices<8, 5, 25>> C1000;
    typedef StockhamCfg<1024, 64, 4, 4, false, Radices<16, 8, 8>> C1024;
    typedef StockhamCfg<2048, 256, 1, 4, false, Radices<8, 8, 8, 4>> C2048;
    typedef StockhamCfg<4096, 256, 1, 4, false, Radices<8, 8, 8, 8>> C4096;
    typedef StockhamCfg<6561, 256, 1, 30, false, Radices<9, 27, 27>, 1, 0, 1, float2, false, true> C6561;
};
constexpr int kClassSets = 2;

template <class C>
constexpr int lds_of() {  // in complex64 units
    return C::SPLITX ? (C::TPB * C::LDS_N + 1) / 2 : C::TPB * C::LDS_N;
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
template <class CS>
constexpr int batch_lds() {
    return cmax(cmax(cmax(lds_of<typename CS::C60>(), lds_of<typename CS::C128>()),
                     cmax(lds_of<typename CS::C243>(), lds_of<typename CS::C1000>())),
                cmax(cmax(cmax(lds_of<typename CS::C1024>(), lds_of<typename CS::C2048>()),
                          cmax(lds_of<typename CS::C4096>(), lds_of<typename CS::C6561>())),
                     cmax(lds_of<typename CS::C512>(), lds_of<typename CS::C256>())));
}
template <class CS>
constexpr bool same_block() {
    return CS::C60::BLOCK == kBatchBlock && CS::C128::BLOCK == kBatchBlock && CS::C243::BLOCK == kBatchBlock &&
           CS::C1000::BLOCK == kBatchBlock && CS::C1024::BLOCK == kBatchBlock && CS::C2048::BLOCK == kBatchBlock &&
           CS::C4096::BLOCK == kBatchBlock && CS::C6561::BLOCK == kBatchBlock && CS::C512::BLOCK == kBatchBlock &&
           CS::C256::BLOCK == kBatchBlock;
}
static_assert(same_block<ClassSet0>() && same_block<ClassSet1>(), "every class runs with the same workgroup size");

// NT: streaming (non-temporal) HBM accesses
template <bool NT>
struct GLoadB {
    const float2* __restrict__ p;
    float sg;
    __device__ __forceinline__ float2 operator()(int i) const {
        float2 v;
        if constexpr (NT)
            v = ld_nt(p + i);
        else
            v = p[i];
        return make_float2(v.x, v.y * sg);
    }
};
template <bool NT>
struct GStoreB {
    float2* __restrict__ p;
    float sc;   // output scale
    float sgs;  // sign * scale
    bool ok;
    __device__ __forceinline__ void operator()(int i, float2 v) const {
        if (ok) {
            if constexpr (NT)
                st_nt(p + i, make_float2(v.x * sc, v.y * sgs));
            else
                p[i] = make_float2(v.x * sc, v.y * sgs);
        }
    }
};

// One tile of class C: transforms [0, tile.count) starting at element offset tile.offset.  Classes
// whose transforms fit in a wave skip empty slots; the others synchronise the whole workgroup
// inside the passes, so every thread runs them and only the stores of empty slots are dropped.
//
// Register budget: the class bodies are inlined into the kernel's dispatch switch, so the kernel
// needs the largest class's registers, not the sum.  Two things keep it that way: the lane index
// is re-read through an opaque move at the top of each body (otherwise the compiler hoists every
// class's per-lane address arithmetic above the switch, where it all stays live: ~200 VGPRs), and
// the loop has a memory barrier per tile (twiddle loads are not hoisted out of the loop).  Calls
// instead of inlining cost more: call arguments travel in VGPRs and the callee-saved register
// ranges of the AMDGPU calling convention inflate each body's footprint.
template <class C, bool LNT, bool SNT>
__device__ __forceinline__ void tile_body(float2* lds, int tl, int tid, bool ok, long long base, const float2* __restrict__ in,
                                          float2* __restrict__ out, const float2* __restrict__ tw, float sg, float sc) {
    GLoadB<LNT> gin{in + base, sg};
    GStoreB<SNT> gout{out + base, sc, sg * sc, ok};
    if constexpr (C::SPLITX)
        split_fft<C>(reinterpret_cast<float*>(lds) + tl * C::LDS_N, tid, tw, gin, gout);
    else
        stockham_passes<C, false, false>(lds + tl * C::LDS_N, tid, tw, gin, gout);
}

// STREAM: streaming (non-temporal) HBM accesses for the classes whose transforms are whole
// 128-byte lines (n a multiple of 16): for plans whose working set is well beyond the last-level
// cache (uneven batch of 27 M points: 89.8 vs 99.8 us).  The odd classes (60, 243, 1000, 6561) keep
// cached accesses: their passes' wave accesses straddle lines, and streaming them measured slower
// (loads 94.6 us, stores 98.1 us, both 102.4 us).
template <class C, bool STREAM>
__device__ __forceinline__ void run_tile(float2* lds, const ofx_batch_tile& tile, const float2* __restrict__ in,
                                         float2* __restrict__ out, const float2* __restrict__ tw, float sg, float sc) {
    int tx;
    __asm__ volatile("v_mov_b32 %0, %1" : "=v"(tx) : "v"((int)threadIdx.x));
    const int tl = tx / C::T;
    const int tid = tx - tl * C::T;
    const bool ok = tl < tile.count;
    if constexpr (C::WAVE_SYNC) {
        if (!ok) return;
    }
    const long long base = tile.offset + (long long)(ok ? tl : 0) * C::N;
    constexpr bool NT = STREAM && C::N % 16 == 0;
    GLoadB<NT> gin{in + base, sg};
    GStoreB<NT> gout{out + base, sc, sg * sc, ok};
    if constexpr (C::SPLITX)
        split_fft<C>(reinterpret_cast<float*>(lds) + tl * C::LDS_N, tid, tw, gin, gout);
    else
        stockham_passes<C, false, false>(lds + tl * C::LDS_N, tid, tw, gin, gout);
}

__global__ __launch_bounds__(kBatchBlock) void k_batch_reset(unsigned int* counter) { *counter = 0u; }

template <class CS, bool STREAM>
__device__ __forceinline__ void run_class(float2* lds, const ofx_batch_tile& tile, const ofx_batch_args& a) {
    const float2* __restrict__ in = static_cast<const float2*>(a.in);
    float2* __restrict__ out = static_cast<float2*>(a.out);
    const float2* tw = static_cast<const float2*>(a.tw) + a.tw_off[tile.cls];
    const float sc = a.scale != 0.f ? a.scale : 1.f;
    switch (tile.cls) {
        case OFX_BCLASS_60: run_tile<typename CS::C60, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_128: run_tile<typename CS::C128, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_243: run_tile<typename CS::C243, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_256: run_tile<typename CS::C256, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_512: run_tile<typename CS::C512, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_1000: run_tile<typename CS::C1000, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_1024: run_tile<typename CS::C1024, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_2048: run_tile<typename CS::C2048, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_4096: run_tile<typename CS::C4096, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        case OFX_BCLASS_6561: run_tile<typename CS::C6561, STREAM>(lds, tile, in, out, tw, a.sign, sc); break;
        default: break;
    }
}

// Tile descriptor `i` (uniform over the workgroup) through the scalar memory path: the list is
// read-only while the kernel runs, and a scalar load is waited for with the LDS counter, so the
// prefetched descriptor does not make the end-of-tile barrier wait for the tile's global stores.
__device__ __forceinline__ ofx_batch_tile load_tile(const ofx_batch_tile* tiles, int i) {
    typedef const __attribute__((address_space(4))) ofx_batch_tile* ConstTiles;
    const ConstTiles p = ((ConstTiles)tiles) + __builtin_amdgcn_readfirstlane(i);
    ofx_batch_tile t;
    t.cls = p->cls;
    t.count = p->count;
    t.offset = p->offset;
    return t;
}

// Position q of a workgroup's walk -> tile index.  The list is sorted by decreasing cost; the first
// P positions take the P smallest tiles in increasing cost (the list's end, backwards), the rest
// the list from its start.  A prologue of small, uneven tiles lets the workgroups drift apart
// before the long tiles come, so the long tiles do not all load, compute and store in lock step.
__device__ __forceinline__ int tile_of(int q, int ntiles, int P) { return q < P ? ntiles - 1 - q : q - P; }

// Every loop below is uniform over the workgroup (tile indices come from blockIdx, the tile list
// or LDS), so every wave leaves together once the index passes ntiles.
template <class CS, int SCHED, bool STREAM>
__global__ __launch_bounds__(kBatchBlock) void k_batch_persistent(ofx_batch_args a) {
    __shared__ float2 lds[batch_lds<CS>()];
    __shared__ int s_next[2];
    const int G = (int)gridDim.x;
    const int b = (int)blockIdx.x;
    // prologue: prologue_rounds tiles per workgroup, at most half the list
    const int P = (int)min((long long)a.prologue_rounds * G, (long long)(a.ntiles / 2));
    if constexpr (SCHED == 1) {
        int t = b;
        ofx_batch_tile cur{};
        if (t < a.ntiles) cur = load_tile(a.tiles, tile_of(t, a.ntiles, P));
        for (int k = 1; t < a.ntiles; ++k) {
            const int tn = k * G + ((k & 1) ? G - 1 - b : b);
            ofx_batch_tile nxt{};
            if (tn < a.ntiles) nxt = load_tile(a.tiles, tile_of(tn, a.ntiles, P));  // in flight while this tile runs
            __asm__ volatile("" ::: "memory");      // no load of the tile bodies moves out of the loop
            run_class<CS, STREAM>(lds, cur, a);
            lds_barrier();  // LDS is reused by the next tile
            cur = nxt;
            t = tn;
        }
    } else {
        int t = b;
        int parity = 0;
        while (t < a.ntiles) {
            unsigned int nx = 0;
            if (threadIdx.x == 0) nx = atomicAdd(a.counter, 1u);  // the next index, used after this tile
            const ofx_batch_tile cur = load_tile(a.tiles, tile_of(t, a.ntiles, P));
            __asm__ volatile("" ::: "memory");
            run_class<CS, STREAM>(lds, cur, a);
            // two slots: a wave may still read this tile's slot while thread 0 writes the next one
            if (threadIdx.x == 0) s_next[parity] = (int)nx + G;
            lds_barrier();  // publishes s_next; LDS is reused by the next tile
            t = s_next[parity];
            parity ^= 1;
        }
    }
}

template <class C>
void fill_info(ofx_batch_class_info* c) {
    c->n = C::N;
    c->transforms_per_tile = C::TPB;
    c->npass = C::NP;
    for (int p = 0; p < C::NP; ++p) c->radices[p] = C::radix(p);
    c->tw_len = C::tw_len();
}

}  // namespace ofx

using namespace ofx;

extern "C" int ofx_batch_class_count(void) { return OFX_BCLASS_COUNT; }
extern "C" int ofx_batch_class_set_count(void) { return kClassSets; }

template <class CS>
static int class_info(int cls, ofx_batch_class_info* c) {
    switch (cls) {
        case OFX_BCLASS_60: fill_info<typename CS::C60>(c); break;
        case OFX_BCLASS_128: fill_info<typename CS::C128>(c); break;
        case OFX_BCLASS_243: fill_info<typename CS::C243>(c); break;
        case OFX_BCLASS_256: fill_info<typename CS::C256>(c); break;
        case OFX_BCLASS_512: fill_info<typename CS::C512>(c); break;
        case OFX_BCLASS_1000: fill_info<typename CS::C1000>(c); break;
        case OFX_BCLASS_1024: fill_info<typename CS::C1024>(c); break;
        case OFX_BCLASS_2048: fill_info<typename CS::C2048>(c); break;
        case OFX_BCLASS_4096: fill_info<typename CS::C4096>(c); break;
        case OFX_BCLASS_6561: fill_info<typename CS::C6561>(c); break;
        default: return (int)hipErrorInvalidValue;
    }
    return 0;
}

extern "C" int ofx_batch_class_info_get_set(int set, int cls, ofx_batch_class_info* c) {
    if (c == nullptr) return (int)hipErrorInvalidValue;
    if (set == 0) return class_info<ClassSet0>(cls, c);
    if (set == 1) return class_info<ClassSet1>(cls, c);
    return (int)hipErrorInvalidValue;
}

extern "C" int ofx_batch_class_info_get(int cls, ofx_batch_class_info* c) { return ofx_batch_class_info_get_set(0, cls, c); }

typedef void (*batch_kernel_t)(ofx_batch_args);
template <class CS>
static batch_kernel_t batch_kernel_of(int schedule, int stream) {
    if (stream)
        return schedule == 0 ? k_batch_persistent<CS, 0, true> : k_batch_persistent<CS, 1, true>;
    return schedule == 0 ? k_batch_persistent<CS, 0, false> : k_batch_persistent<CS, 1, false>;
}
static batch_kernel_t batch_kernel(int set, int schedule, int stream) {
    if (set == 0) return batch_kernel_of<ClassSet0>(schedule, stream);
    if (set == 1) return batch_kernel_of<ClassSet1>(schedule, stream);
    return nullptr;
}

extern "C" int ofx_batch_blocks_per_cu_stream(int set, int schedule, int stream) {
    // cached per device, class set, schedule and access kind (the query is a runtime call; plans
    // ask on every execution, possibly from several host threads)
    static std::atomic<int> cache[64][kClassSets][2][2] = {};
    if (set < 0 || set >= kClassSets) return 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = -1;
    const int sc = schedule == 0 ? 0 : 1;
    const int st = stream ? 1 : 0;
    if (dev >= 0) {
        const int c = cache[dev][set][sc][st].load(std::memory_order_relaxed);
        if (c > 0) return c;
    }
    int n = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel(set, sc, st), kBatchBlock, 0);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        n = 1;
    }
    if (dev >= 0) cache[dev][set][sc][st].store(n, std::memory_order_relaxed);
    return n;
}

extern "C" int ofx_batch_blocks_per_cu_set(int set, int schedule) { return ofx_batch_blocks_per_cu_stream(set, schedule, 0); }

extern "C" int ofx_batch_blocks_per_cu(int schedule) { return ofx_batch_blocks_per_cu_set(0, schedule); }

extern "C" int ofx_launch_batch_persistent(const ofx_batch_args* a, int blocks, hipStream_t stream) {
    if (a == nullptr || a->tiles == nullptr || blocks <= 0 || (a->schedule == 0 && a->counter == nullptr) ||
        a->class_set < 0 || a->class_set >= kClassSets || (a->schedule != 0 && a->schedule != 1) ||
        (a->stream != 0 && a->stream != 1))
        return (int)hipErrorInvalidValue;
    if (a->ntiles <= 0) return 0;
    if (a->schedule == 0) {
        hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, stream, a->counter);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    const batch_kernel_t k = batch_kernel(a->class_set, a->schedule, a->stream);
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kBatchBlock), 0, stream, *a);
    return (int)hipGetLastError();
}
