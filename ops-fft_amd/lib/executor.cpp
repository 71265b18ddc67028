// executor.cpp - kernel choice, device tables, staging and launch sequences of a plan.
#include "executor.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <map>
#include <sstream>

#include "tune_cache.h"

namespace ofx {

namespace {

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

bool is_device_pointer(const void* p) {
    hipPointerAttribute_t attr;
    std::memset(&attr, 0, sizeof attr);
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // clear the sticky error of an unknown (pageable host) pointer
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified;
}

const std::vector<ofx_kernel_info>& registry() {
    static std::vector<ofx_kernel_info> reg = [] {
        std::vector<ofx_kernel_info> r;
        const int n = ofx_kernel_count();
        r.resize(n);
        for (int i = 0; i < n; ++i) ofx_kernel_info_get(i, &r[i]);
        return r;
    }();
    return reg;
}

// Report a failed HIP call on stderr (the library's "[ops-fft]" convention) and pass the code on.
hipError_t hip_log(hipError_t e, const char* what) {
    if (e != hipSuccess) std::fprintf(stderr, "[ops-fft] %s failed: %s (%d)\n", what, hipGetErrorString(e), (int)e);
    return e;
}

const std::vector<ofx_2d_kernel_info>& reg_2d() {
    static std::vector<ofx_2d_kernel_info> reg = [] {
        std::vector<ofx_2d_kernel_info> r(ofx_2d_kernel_count());
        for (int i = 0; i < (int)r.size(); ++i) ofx_2d_kernel_info_get(i, &r[i]);
        return r;
    }();
    return reg;
}

const std::vector<ofx_kernel_info>& col_registry() {
    static std::vector<ofx_kernel_info> reg = [] {
        std::vector<ofx_kernel_info> r(ofx_col_kernel_count());
        for (int i = 0; i < (int)r.size(); ++i) ofx_col_kernel_info_get(i, &r[i]);
        return r;
    }();
    return reg;
}

int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

}  // namespace

// -------------------------------------------------------------------------------------------
// plan construction
// -------------------------------------------------------------------------------------------

// Radix sequence for the runtime-radix kernel: prime factors grouped into supported codelets
// (2^k -> 16/8/4/2, 3^k -> 9/3), then the two smallest radices merged while their product is a
// supported codelet of at most 16 (e.g. 3000 -> 8 x 15 x 25 is not reached; 3000 -> 8,3,5,5,5 ->
// 8,15,5,5 -> 8,15,25).  Empty result: a prime factor above 47.
std::vector<int> generic_radices(long long n) {
    std::vector<int> f;
    if (n <= 1) return f;
    long long m = n;
    int twos = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++twos;
    }
    while (twos >= 4) {
        f.push_back(16);
        twos -= 4;
    }
    if (twos > 0) f.push_back(1 << twos);
    int threes = 0;
    while (m % 3 == 0) {
        m /= 3;
        ++threes;
    }
    while (threes >= 2) {
        f.push_back(9);
        threes -= 2;
    }
    if (threes) f.push_back(3);
    for (long long p = 5; p * p <= m || m > 1; p += 2) {
        if (p * p > m) p = m;
        while (m % p == 0) {
            if (!ofx_generic_radix_supported((int)p)) return {};
            f.push_back((int)p);
            m /= p;
        }
        if (m == 1) break;
    }
    for (;;) {
        std::sort(f.begin(), f.end());
        if (f.size() < 2) break;
        const int prod = f[0] * f[1];
        if (prod > 25 || !ofx_generic_radix_supported(prod)) break;
        f.erase(f.begin(), f.begin() + 2);
        f.push_back(prod);
    }
    std::sort(f.begin(), f.end(), [](int a, int b) { return a > b; });
    return f;
}

struct Builder {
    Executor* e;

    // All registry entries computing `kind` at complex length n, in preference order; staged
    // entries only when the transforms are contiguous.
    // One-pass 2D C2C (complex64) of `images` nx x ny images when a compiled LDS-resident kernel
    // covers the size (and ACLFFT_2D_FUSED is not 0); false otherwise.
    bool fused_2d(long long nx, long long ny, long long images) {
        if (env_int("ACLFFT_2D_FUSED", 1) == 0 || env_int("ACLFFT_DISABLE_TEMPLATED", 0) != 0) return false;
        const auto& reg = reg_2d();
        for (int i = 0; i < (int)reg.size(); ++i) {
            if (reg[i].nx != nx || reg[i].ny != ny) continue;
            Step s;
            s.type = Step::FUSED2D;
            s.kernel = i;
            s.n = (int)(nx * ny);
            s.tw_table = pass_table_for(reg[i].row_npass, reg[i].row_radices, reg[i].row_tw_len);
            s.aux_table = pass_table_for(reg[i].col_npass, reg[i].col_radices, reg[i].col_tw_len);
            s.args.batch = images;
            s.args.idist = s.args.odist = nx * ny;
            s.args.in_sign = s.args.out_sign = s.args.scale = 1.f;
            // the pipelined form (next image's loads in flight) for the sizes it covers, unless
            // ACLFFT_2D_PIPE=0
            if (env_int("ACLFFT_2D_PIPE", 1) == 0) s.args.flags |= OFX_FLAG_NO_PIPELINE;
            s.conj_in = s.conj_out = true;
            e->steps_.push_back(s);
            return true;
        }
        return false;
    }

    std::vector<int> kernels_for(int kind, long long n, bool contiguous) {
        std::vector<int> ids;
        if (env_int("ACLFFT_DISABLE_TEMPLATED", 0) != 0) return ids;
        const auto& reg = registry();
        for (int i = 0; i < (int)reg.size(); ++i) {
            if (reg[i].kind == kind && reg[i].n == n && (contiguous || !reg[i].staged)) ids.push_back(i);
        }
        return ids;
    }

    int pass_table_for(int npass, const int* radices, long long len) {
        if (len <= 0) return -1;
        TableSpec t;
        t.kind = TableSpec::PASS;
        t.npass = npass;
        for (int p = 0; p < npass; ++p) t.radices[p] = radices[p];
        t.bytes = (size_t)len * 8;
        return e->add_table(t);
    }
    int pass_table_z(int npass, const int* radices, long long len) {
        if (len <= 0) return -1;
        TableSpec t;
        t.kind = TableSpec::PASS_Z;
        t.npass = npass;
        for (int p = 0; p < npass; ++p) t.radices[p] = radices[p];
        t.bytes = (size_t)len * 16;
        return e->add_table(t);
    }
    int pass_table(const ofx_kernel_info& k) {
        if (!is_z_kind(k.kind)) return pass_table_for(k.npass, k.radices, k.tw_len);
        // complex128 kernel: the table in FP64 (16-byte entries)
        if (k.tw_len <= 0) return -1;
        TableSpec t;
        t.kind = TableSpec::PASS_Z;
        t.npass = k.npass;
        for (int p = 0; p < k.npass; ++p) t.radices[p] = k.radices[p];
        t.bytes = (size_t)k.tw_len * 16;
        return e->add_table(t);
    }
    static bool is_z_kind(int k) { return k >= OFX_KIND_Z2Z && k <= OFX_KIND_COL_Z; }
    int roots_table_z(long long len, long long count) {
        TableSpec t;
        t.kind = TableSpec::ROOTS_Z;
        t.len = len;
        t.count = count;
        t.bytes = (size_t)count * 16;
        return e->add_table(t);
    }
    int roots_table(long long len, long long count) {
        TableSpec t;
        t.kind = TableSpec::ROOTS;
        t.len = len;
        t.count = count;
        t.bytes = (size_t)count * 8;
        return e->add_table(t);
    }
    int fourstep_table(long long len) {
        TableSpec t;
        t.kind = TableSpec::FOURSTEP;
        t.len = len;
        t.count = ofx_ft_table_count(len);
        t.bytes = (size_t)t.count * 16;
        return e->add_table(t);
    }

    Step kernel_step(int kind, long long n, const std::vector<int>& ids) {
        Step s;
        s.type = Step::KERNEL;
        s.kind = kind;
        s.n = (int)n;
        s.candidates = ids;
        for (int id : ids) s.candidate_tables.push_back(pass_table(registry()[id]));
        const int pick = std::min<int>(std::max(0, env_int("ACLFFT_KERNEL_VARIANT", 0)), (int)ids.size() - 1);
        s.kernel = ids[pick];
        s.tw_table = s.candidate_tables[pick];
        s.args.in_sign = 1.f;
        s.args.out_sign = 1.f;
        s.args.scale = 1.f;
        return s;
    }

    // Streaming kernels (non-temporal HBM accesses) pay off when a launch's input plus output is
    // well beyond the 256 MB last-level cache; below that, repeated or chained launches find part
    // of their data in that cache, which streaming accesses forgo.  Unless a variant is forced
    // (ACLFFT_KERNEL_VARIANT), the step takes the first candidate of the wanted kind.
    static constexpr size_t kStreamBytes = size_t(384) << 20;
    void pick_stream(Step& s, size_t io_bytes) {
        if (std::getenv("ACLFFT_KERNEL_VARIANT") != nullptr) return;
        const bool want = io_bytes >= kStreamBytes;
        for (size_t c = 0; c < s.candidates.size(); ++c)
            if ((registry()[s.candidates[c]].stream != 0) == want) {
                s.kernel = s.candidates[c];
                s.tw_table = s.candidate_tables[c];
                return;
            }
    }

    // Store rule of single-launch plans: when the plan is one kernel step writing at least
    // 16 MiB of output and below the streaming threshold (pick_stream), the step takes its first
    // candidate with streaming stores and cached loads (StockhamCfg::STREAM 2): this plan never
    // re-reads its output, and write-allocating it in the last-level cache evicts the input and
    // forces write-backs.  Measured on MI355X: C2C 4096 x 1024 10.7 vs 13.8 us, C2C 2048 x 8192
    // 39.6 vs 41.8 us, R2C / C2R 8192 x 1024 13.0 vs 15.5 / 15.9 us, C2R 4096 x 8192 43.6 vs
    // 46.2 us, R2C / C2R 8192 x 8192 94.0 / 98.1 vs 115.7 / 124.1 us.  A consumer that reads the
    // output right after the call then reads it from HBM, not from the cache
    // (ACLFFT_STREAM_STORES=0 keeps cached stores).
    //
    // R2C plans stream their stores only below 64 MiB of output: from a cold last-level cache the
    // larger ones run faster with cached stores (4096 x 8192: 56.8 vs 62.4 us, 3000 x 8192: 44.3 vs
    // 45.6 us), and back-to-back calls on one input tie (43.7-45.2 vs 44.1-44.7 us, 33.0-34.4 vs
    // 33.6 us); 8192 x 1024 (34 MB) gains in both (cold 18.5 vs 19.0 us).
    static constexpr size_t kStoreStreamBytes = size_t(16) << 20;
    static constexpr size_t kR2CStoreStreamMax = size_t(64) << 20;
    void stream_stores(size_t out_bytes) {
        if (e->steps_.size() != 1 || out_bytes < kStoreStreamBytes) return;
        if (std::getenv("ACLFFT_KERNEL_VARIANT") != nullptr || env_int("ACLFFT_STREAM_STORES", 1) == 0) return;
        Step& s = e->steps_[0];
        if (s.type != Step::KERNEL || s.dst != BUF_OUT || registry()[s.kernel].stream != 0) return;
        if (registry()[s.kernel].kind == OFX_KIND_R2C && out_bytes >= kR2CStoreStreamMax) return;
        for (size_t c = 0; c < s.candidates.size(); ++c)
            if (registry()[s.candidates[c]].stream == 2) {
                s.kernel = s.candidates[c];
                s.tw_table = s.candidate_tables[c];
                return;
            }
    }

    // The same rule for 2D plans that run as the one-pass kernel: streaming stores for >= 16 MiB
    // of output, and streaming loads as well when input plus output is >= 384 MiB (as the 1D
    // kernels do).  Measured at 2^26 points: 32^2 173.5 vs 181.3 us, 64^2 200.0 vs 205.3 us,
    // 128^2 259.8 vs 286.0 us (2^22 points of 128^2: 14.9 vs 17.8 us).  Column passes keep cached
    // stores: their groups write 64-128-byte runs of a row, and streaming them measured slower
    // (2048^2 x 4 102 vs 92 us, 1024^2 x 64 410 vs 389 us).
    void stream_final_multi(size_t in_bytes, size_t out_bytes) {
        if (e->steps_.size() != 1 || env_int("ACLFFT_STREAM_STORES", 1) == 0) return;
        Step& s = e->steps_.back();
        if (s.type != Step::FUSED2D || s.dst != BUF_OUT) return;
        if (out_bytes >= kStoreStreamBytes) s.args.flags |= OFX_FLAG_NT_STORE;
        if (in_bytes + out_bytes >= kStreamBytes) s.args.flags |= OFX_FLAG_NT_LOAD;
    }

    // A runtime-radix step of length n; false if n has an unsupported prime factor or is too long.
    static long long generic_max(bool z) { return z ? OFX_GENERIC_MAX_ELEMS_Z : OFX_GENERIC_MAX_ELEMS; }
    bool generic_step(long long n, Step* out, bool z = false) {
        if (n > generic_max(z)) return false;
        std::vector<int> r = generic_radices(n);
        if ((r.empty() && n != 1) || (int)r.size() > OFX_MAX_PASSES) return false;
        Step s;
        s.type = Step::GENERIC;
        s.n = (int)n;
        ofx_generic_desc& g = s.g;
        std::memset(&g, 0, sizeof g);
        g.n = (int)n;
        g.npass = (int)r.size();
        for (int p = 0; p < g.npass; ++p) g.radices[p] = r[p];
        // pass p >= 1 starts after the entries of passes 1..p-1 (pass 0 needs no twiddles)
        long long acc = 0, ns2 = r.empty() ? 1 : r[0];
        g.tw_off[0] = 0;
        for (int p = 1; p < g.npass; ++p) {
            g.tw_off[p] = (int)acc;
            acc += ns2 * (r[p] - 1);
            ns2 *= r[p];
        }
        g.z = z ? 1 : 0;
        s.tw_table = z ? pass_table_z(g.npass, g.radices, acc) : pass_table_for(g.npass, g.radices, acc);
        // points per workgroup: enough bytes in flight per CU (ACLFFT_GENERIC_ELEMS overrides)
        const long long per_block = std::max<long long>(
            64, std::min<long long>(generic_max(z), env_int("ACLFFT_GENERIC_ELEMS", z ? 1024 : 2048)));
        g.tpb = (int)std::max<long long>(1, std::min<long long>(64, per_block / n));
        g.cnt_lo = 1;
        s.args.in_sign = 1.f;
        s.args.out_sign = 1.f;
        s.args.scale = 1.f;
        *out = s;
        return true;
    }

    // Batch of contiguous C2C transforms of length n from buffer src to dst.  The direction's
    // conjugation is applied on the first load (conj_in) and the last store (conj_out).
    aclfftResult c2c_rows(long long n, long long batch, int src, int dst, bool conj_in, bool conj_out, bool z = false) {
        // (a copy cannot scale: with a scale, length 1 runs as the pass-free runtime-radix step)
        if (n == 1 && e->desc_.scale == 1.f) {
            Step s;
            s.type = Step::COPY;
            s.src = src;
            s.dst = dst;
            s.copy_bytes = (size_t)batch * (z ? 16 : 8);
            e->steps_.push_back(s);
            // a length-1 transform is the identity in both directions
            return ACLFFT_SUCCESS;
        }
        const int kind = z ? OFX_KIND_Z2Z : OFX_KIND_C2C;
        auto ids = kernels_for(kind, n, true);
        if (!ids.empty()) {
            Step s = kernel_step(kind, n, ids);
            pick_stream(s, (size_t)n * batch * (z ? 32 : 16));
            s.src = src;
            s.dst = dst;
            s.args.batch = batch;
            s.args.idist = n;
            s.args.odist = n;
            s.conj_in = conj_in;
            s.conj_out = conj_out;
            e->steps_.push_back(s);
            return ACLFFT_SUCCESS;
        }
        Step s;
        if (generic_step(n, &s, z)) {
            s.src = src;
            s.dst = dst;
            s.args.batch = batch;
            s.args.idist = n;
            s.args.odist = n;
            s.args.istride = 1;
            s.args.ostride = 1;
            s.conj_in = conj_in;
            s.conj_out = conj_out;
            e->steps_.push_back(s);
            return ACLFFT_SUCCESS;
        }
        return c2c_fourstep(n, batch, src, dst, conj_in, conj_out, z);
    }

    // Split n = n1 * n2 for the four-step: both factors runnable by the runtime-radix kernel, as
    // close to sqrt(n) as possible with n1 (the strided column length) the smaller one.
    bool split(long long n, long long* n1, long long* n2, bool z = false) {
        long long best = 0;
        for (long long a = 2; a * a <= n; ++a) {
            if (n % a) continue;
            const long long b = n / a;
            if (a > generic_max(z) || b > generic_max(z)) continue;
            if (generic_radices(a).empty() || generic_radices(b).empty()) continue;
            best = a;
        }
        if (best == 0) return false;
        *n1 = best;
        *n2 = n / best;
        return true;
    }

    // A four-step split whose column length has a compiled column kernel and whose row length has a
    // compiled transposed-store row kernel, as balanced as possible.
    static int col_kind(bool z) { return z ? OFX_KIND_COL_Z : OFX_KIND_COL; }
    static int rowt_kind(bool z) { return z ? OFX_KIND_ROWT_Z : OFX_KIND_ROWT; }
    bool templated_split(long long n, long long* n1, long long* n2, bool z = false) {
        if (env_int("ACLFFT_DISABLE_TEMPLATED", 0) != 0) return false;
        double best = 1e30;
        for (const ofx_kernel_info& c : col_registry()) {
            if (c.kind != col_kind(z) || n % c.n) continue;
            const long long b = n / c.n;
            if (kernels_for(rowt_kind(z), b, true).empty()) continue;
            const double d = std::fabs(std::log((double)c.n) - std::log((double)b));
            if (d < best) {
                best = d;
                *n1 = c.n;
                *n2 = b;
            }
        }
        return best < 1e30;
    }

    // A three-factor split n = n1 * n2 * n3 with compiled column kernels for n1 and n2 and a
    // transposed-store row kernel for n3, minimising the largest factor.
    bool templated_split3(long long n, long long* n1, long long* n2, long long* n3, bool z = false) {
        if (env_int("ACLFFT_DISABLE_TEMPLATED", 0) != 0) return false;
        long long best = 0;
        for (const ofx_kernel_info& a : col_registry()) {
            if (a.kind != col_kind(z) || n % a.n) continue;
            for (const ofx_kernel_info& b : col_registry()) {
                if (b.kind != col_kind(z) || (n / a.n) % b.n) continue;
                const long long c = n / a.n / b.n;
                if (kernels_for(rowt_kind(z), c, true).empty()) continue;
                const long long worst = std::max<long long>(std::max(a.n, b.n), c);
                if (best == 0 || worst < best) {
                    best = worst;
                    *n1 = a.n;
                    *n2 = b.n;
                    *n3 = c;
                }
            }
        }
        return best > 0;
    }

    // Four-step twiddles of transforms of n >= 2^20 points are formed in FP64 and rounded once
    // (OFX_FLAG_FT_FP64); shorter ones from FP32 factors (about 2 ulp).  ACLFFT_FOURSTEP_FP64=0|1
    // forces either form.
    static void ft_precision(Step& st, long long n) {
        const int f = env_int("ACLFFT_FOURSTEP_FP64", -1);
        if (f == 1 || (f < 0 && n >= (1LL << 20))) st.args.flags |= OFX_FLAG_FT_FP64;
    }

    // Three-level four-step, n = N1*N2*N3, index n = n1*N2*N3 + n2*N3 + n3, k = k1 + N1*k2 + N1*N2*k3:
    //   A: columns over n1 (stride N2*N3), twiddle W_n^(k1*(n2*N3 + n3)), src -> dst (in place safe)
    //   B: per (b, k1), columns over n2 (stride N3), twiddle W_(N2*N3)^(k2*n3), stored to the
    //      workspace as [k2][b][k1][n3] so that the rows of step C that land next to each other
    //      in X are next to each other in the workspace
    //   C: rows over n3, transposed store X[b*n + k1 + N1*k2 + N1*N2*k3].
    aclfftResult c2c_fourstep3(long long n, long long batch, long long N1, long long N2, long long N3, int src,
                               int dst, bool conj_in, bool conj_out, bool z = false) {
        const int scratch = (dst == BUF_WORK0) ? BUF_WORK1 : BUF_WORK0;
        const size_t elem = z ? 16 : 8;
        aclfftResult r = c2c_columns(N1, N2 * N3, batch, src, dst, conj_in, false, z);
        if (r != ACLFFT_SUCCESS) return r;
        e->steps_.back().args.ft_len = n;
        e->steps_.back().aux_table = fourstep_table(n);
        ft_precision(e->steps_.back(), n);
        r = c2c_columns(N2, N3, batch * N1, dst, scratch, false, false, z);
        if (r != ACLFFT_SUCCESS) return r;
        Step& b = e->steps_.back();
        b.args.odist = N3;
        b.args.ostride = batch * N1 * N3;
        b.args.ft_len = N2 * N3;
        b.aux_table = fourstep_table(N2 * N3);
        ft_precision(b, n);
        Step c = kernel_step(rowt_kind(z), N3, kernels_for(rowt_kind(z), N3, true));
        c.src = scratch;
        c.dst = dst;
        c.args.batch = batch * N1 * N2;
        c.args.idist = N3;
        c.args.cols = batch * N1;
        c.args.cols_lo = N1;
        c.args.odist = N1;
        c.args.odist_hi = n;
        c.args.ostride = N1 * N2;
        c.conj_out = conj_out;
        e->steps_.push_back(c);
        size_t& wb = (scratch == BUF_WORK0) ? e->work_bytes_ : e->work1_bytes_;
        wb = std::max(wb, (size_t)(n * batch) * elem);
        return ACLFFT_SUCCESS;
    }

    // Four-step: x[n1*N2 + n2] -> columns (length N1, twiddle W_N^(n2 k1)) -> workspace
    // tmp[k1*N2 + n2] -> rows (length N2) -> X[k1 + N1*k2].  Two factors when both are at most
    // 2048 (whole columns and rows in one workgroup with 64-128-byte HBM runs), else three
    // factors; ACLFFT_FOURSTEP_LEVELS=2|3 forces a form where it exists.
    aclfftResult c2c_fourstep(long long n, long long batch, int src, int dst, bool conj_in, bool conj_out,
                              bool z = false) {
        long long T1 = 0, T2 = 0, U1 = 0, U2 = 0, U3 = 0;
        const size_t elem = z ? 16 : 8;
        const bool two = templated_split(n, &T1, &T2, z);
        const bool three = templated_split3(n, &U1, &U2, &U3, z);
        const int force = env_int("ACLFFT_FOURSTEP_LEVELS", 0);
        // two factors up to 2048 x 2048 (measured on MI355X, 2^22 x 16: 572 us as 2048 x 2048 vs
        // 730 us as 64 x 256 x 256); three beyond (2^24 x 4: 664 us vs 931 us as 4096 x 4096)
        bool use3 = three && (!two || std::max(T1, T2) > 2048);
        if (force == 2 && two) use3 = false;
        if (force == 3 && three) use3 = true;
        if (use3) return c2c_fourstep3(n, batch, U1, U2, U3, src, dst, conj_in, conj_out, z);
        if (two) {
            // The workspace rows are P = N2 + pad elements apart: with a power-of-two pitch every
            // row of a column group starts on the same HBM channel set.  Measured on MI355X: a
            // 16-element pad gains ~2% for N2 >= 512 and loses 1-2% for shorter rows
            // (ACLFFT_FOURSTEP_PAD overrides it).
            const long long P = T2 + std::max(0, env_int("ACLFFT_FOURSTEP_PAD", T2 >= 512 ? 16 : 0));
            // column step: N2 columns of length N1 per transform, four-step twiddle fused in the store
            aclfftResult r = c2c_columns(T1, T2, batch, src, BUF_WORK0, conj_in, false, z);
            if (r != ACLFFT_SUCCESS) return r;
            Step& a = e->steps_.back();
            a.args.ft_len = n;
            a.args.odist = T1 * P;
            a.args.ostride = P;
            a.aux_table = fourstep_table(n);
            ft_precision(a, n);
            // row step: rows of length N2, transposed store X[k1 + N1*k2]
            auto ids = kernels_for(rowt_kind(z), T2, true);
            Step b = kernel_step(rowt_kind(z), T2, ids);
            b.src = BUF_WORK0;
            b.dst = dst;
            b.args.batch = batch * T1;
            b.args.idist = P;
            b.args.cols = T1;
            b.args.odist = n;
            b.args.ostride = T1;
            b.conj_out = conj_out;
            e->steps_.push_back(b);
            e->work_bytes_ = std::max(e->work_bytes_, (size_t)(batch * T1 * P) * elem);
            return ACLFFT_SUCCESS;
        }
        long long N1 = 0, N2 = 0;
        if (!split(n, &N1, &N2, z)) return ACLFFT_NOT_IMPLEMENTED;
        const int ft = fourstep_table(n);
        Step a, b;
        if (!generic_step(N1, &a, z) || !generic_step(N2, &b, z)) return ACLFFT_NOT_IMPLEMENTED;
        // column step
        a.g.tpb = (int)std::max<long long>(1, std::min<long long>(16, generic_max(z) / N1));
        a.g.cnt_lo = N2;
        a.g.idist_lo = 1;
        a.g.odist_lo = 1;
        a.g.in_tmajor = 1;
        a.g.out_tmajor = 1;
        a.g.ft_len = n;
        a.aux_table = ft;
        a.src = src;
        a.dst = BUF_WORK0;
        a.args.batch = batch * N2;
        a.args.idist = n;
        a.args.odist = n;
        a.args.istride = N2;
        a.args.ostride = N2;
        a.conj_in = conj_in;
        // row step
        b.g.tpb = (int)std::max<long long>(1, std::min<long long>(16, generic_max(z) / N2));
        b.g.cnt_lo = N1;
        b.g.idist_lo = N2;
        b.g.odist_lo = 1;
        b.g.in_tmajor = 0;
        b.g.out_tmajor = 1;
        b.src = BUF_WORK0;
        b.dst = dst;
        b.args.batch = batch * N1;
        b.args.idist = n;
        b.args.odist = n;
        b.args.istride = 1;
        b.args.ostride = N1;
        b.conj_out = conj_out;
        e->steps_.push_back(a);
        e->steps_.push_back(b);
        e->work_bytes_ = std::max(e->work_bytes_, (size_t)(n * batch) * elem);
        return ACLFFT_SUCCESS;
    }

    // Columns too long for one workgroup ([images][n][cols], n = N1 * N2, both factors with
    // compiled column kernels): row r = n1*N2 + n2, output row k1 + N1*k2.
    //   A: per image, the N2*cols interleaved columns of length N1 (column n2*cols + c), twiddle
    //      W_n^(k1*n2) (the column index divided by cols), into the workspace at row k1*N2 + n2
    //   B: per (image, k1), the cols columns of length N2 of workspace rows k1*N2 .. k1*N2 + N2-1,
    //      stored to row k1 + N1*k2 of that image of the output.
    // Returns NOT_IMPLEMENTED, with no step added, when n has no such split.
    aclfftResult column_fourstep(long long n, long long cols, long long images, int src, int dst, bool conj_in,
                                 bool conj_out, bool z) {
        if (env_int("ACLFFT_DISABLE_TEMPLATED", 0) != 0) return ACLFFT_NOT_IMPLEMENTED;
        long long N1 = 0, N2 = 0;
        double best = 1e30;
        for (const ofx_kernel_info& a : col_registry()) {
            if (a.kind != col_kind(z) || n % a.n || a.n == n) continue;
            const long long m = n / a.n;
            bool has = false;
            for (const ofx_kernel_info& b : col_registry()) has = has || (b.kind == col_kind(z) && b.n == m);
            if (!has) continue;
            const double dist = std::fabs(std::log((double)a.n) - std::log((double)m));
            if (dist < best) {
                best = dist;
                N1 = a.n;
                N2 = m;
            }
        }
        if (N1 == 0) return ACLFFT_NOT_IMPLEMENTED;
        const int ws = (dst == BUF_WORK0 || src == BUF_WORK0) ? BUF_WORK1 : BUF_WORK0;
        aclfftResult r = c2c_columns(N1, N2 * cols, images, src, ws, conj_in, false, z);
        if (r != ACLFFT_SUCCESS) return r;
        Step& a = e->steps_.back();
        a.args.ft_len = n;
        a.args.ft_div = cols;
        a.aux_table = fourstep_table(n);
        ft_precision(a, n);
        r = c2c_columns(N2, cols, N1 * images, ws, dst, false, conj_out, z);
        if (r != ACLFFT_SUCCESS) return r;
        Step& b = e->steps_.back();
        b.args.odist = cols;
        b.args.cols_lo = N1;
        b.args.odist_hi = n * cols;
        b.args.ostride = N1 * cols;
        size_t& wb = (ws == BUF_WORK0) ? e->work_bytes_ : e->work1_bytes_;
        wb = std::max(wb, (size_t)(n * cols * images) * (z ? 16 : 8));
        return ACLFFT_SUCCESS;
    }

    // Columns of [hi][rows][cols] row-major data: `cols` transforms of length `rows` per image.
    aclfftResult c2c_columns(long long rows, long long cols, long long images, int src, int dst, bool conj_in,
                             bool conj_out, bool z = false) {
        std::vector<int> ids;
        if (env_int("ACLFFT_DISABLE_TEMPLATED", 0) == 0) {
            const auto& creg = col_registry();
            for (int i = 0; i < (int)creg.size(); ++i)
                if (creg[i].n == rows && creg[i].kind == col_kind(z)) ids.push_back(i);
            // Large column sets: split-exchange variants (half the LDS per group, so two groups
            // share a CU) first.  Measured on MI355X, 1024-point columns: 392 vs 418 us for
            // 64 x 1024 x 1024 (2D), 49.8 vs 55.3 us for 1024 x 16384 (VERTICAL); for 8 MiB
            // (4 x 1024 x 1024) the full-LDS variant is ahead, 27.3 vs 29.7 us.
            const size_t elem = z ? 16 : 8;
            if ((size_t)(rows * cols * images) * elem >= ((size_t)64 << 20)) {
                std::stable_partition(ids.begin(), ids.end(), [&](int i) {
                    return (size_t)creg[i].lds_bytes < (size_t)creg[i].n * creg[i].transforms_per_block * elem;
                });
            }
            // widest column groups first only when a row has enough columns to fill them
            std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) {
                const bool fa = creg[a].transforms_per_block <= cols, fb = creg[b].transforms_per_block <= cols;
                return fa > fb;
            });
        }
        if (!ids.empty()) {
            Step c;
            c.type = Step::COLUMN;
            c.kind = col_kind(z);
            c.n = (int)rows;
            c.candidates = ids;
            for (int id : ids) c.candidate_tables.push_back(pass_table(col_registry()[id]));
            const int pick = std::min<int>(std::max(0, env_int("ACLFFT_COLUMN_VARIANT", 0)), (int)ids.size() - 1);
            c.kernel = ids[pick];
            c.tw_table = c.candidate_tables[pick];
            c.args.in_sign = c.args.out_sign = c.args.scale = 1.f;
            c.args.batch = cols * images;
            c.args.cols = cols;
            c.args.idist = rows * cols;
            c.args.odist = rows * cols;
            c.args.istride = cols;
            c.args.ostride = cols;
            c.src = src;
            c.dst = dst;
            c.conj_in = conj_in;
            c.conj_out = conj_out;
            e->steps_.push_back(c);
            return ACLFFT_SUCCESS;
        }
        // no single column kernel: the column four-step over compiled factors, else the
        // runtime-radix kernel
        if (column_fourstep(rows, cols, images, src, dst, conj_in, conj_out, z) == ACLFFT_SUCCESS)
            return ACLFFT_SUCCESS;
        Step s;
        if (!generic_step(rows, &s, z)) return ACLFFT_NOT_IMPLEMENTED;
        s.g.tpb = (int)std::max<long long>(1, std::min<long long>(16, generic_max(z) / rows));
        s.g.cnt_lo = cols;
        s.g.idist_lo = 1;
        s.g.odist_lo = 1;
        s.g.in_tmajor = 1;
        s.g.out_tmajor = 1;
        s.src = src;
        s.dst = dst;
        s.args.batch = cols * images;
        s.args.idist = rows * cols;
        s.args.odist = rows * cols;
        s.args.istride = cols;
        s.args.ostride = cols;
        s.conj_in = conj_in;
        s.conj_out = conj_out;
        e->steps_.push_back(s);
        return ACLFFT_SUCCESS;
    }

    // MFMA split-precision C2C: n = r1 * r2 as two dense DFT stages (r1, r2 in {16, 32, 64}; the
    // registry's lengths 256 ... 4096), or n = 16, 32, 64 as one dense stage (r1 = 1).  Every registry entry of length n is a candidate (the
    // autotuner times them); the first one is the default (for 1024 the 32 x 32 kernel on
    // v_mfma_f32_32x32x16).  ACLFFT_MFMA_VARIANT picks another candidate.
    static std::vector<int> mfma_kernels_for(long long n) {
        std::vector<int> ids;
        std::vector<int> tiled;
        for (int i = 0; i < ofx_mfma_kernel_count(); ++i) {
            ofx_mfma_info mi;
            if (ofx_mfma_kernel_info_get(i, &mi) == 0 && mi.n == n) ids.push_back(i);
            if (ofx_mfma_kernel_info_get(i, &mi) == 0 && mi.n == n) tiled.push_back(mi.tiled);
        }
        // 4096 (64 x 64): the tiled form first (one row tile of input in registers instead of
        // the whole transform: measured 202 vs 246 us (BF16) and 301 vs 339 us (FP16 hi/lo) for
        // 64 Mi points)
        if (n > 2048) {
            std::vector<int> order(ids.size());
            for (size_t k = 0; k < ids.size(); ++k) order[k] = (int)k;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tiled[a] > tiled[b]; });
            std::vector<int> sorted;
            for (int k : order) sorted.push_back(ids[k]);
            ids = sorted;
        }
        return ids;
    }
    // the MFMA kernels of length n that run tile lists (uneven plans): the two-stage ones other
    // than entry 0 (the dense single-stage kernels take no tile lists)
    static std::vector<int> mfma_tile_kernels_for(long long n) {
        std::vector<int> ids;
        for (int id : mfma_kernels_for(n)) {
            ofx_mfma_info mi;
            if (id != 0 && ofx_mfma_kernel_info_get(id, &mi) == 0 && mi.r1 >= 16) ids.push_back(id);
        }
        return ids;
    }
    int mfma_table(int id, int mode) {
        TableSpec t;
        t.kind = TableSpec::MFMA_CONST;
        t.npass = mode;
        t.radices[0] = id;
        t.bytes = (size_t)ofx_mfma_kernel_const_bytes(id, mode);
        return e->add_table(t);
    }
    aclfftResult c2c_mfma(long long n, long long batch, int mode, int src = BUF_IN, int dst = BUF_OUT,
                          bool conj = true) {
        const std::vector<int> ids = mfma_kernels_for(n);
        if (ids.empty() || ofx_mfma_kernel_const_bytes(ids[0], mode) <= 0) return ACLFFT_NOT_SUPPORTED;
        Step s;
        s.type = Step::MFMA;
        s.n = (int)n;
        s.mfma_mode = mode;
        s.candidates = ids;
        for (int id : ids) s.candidate_tables.push_back(mfma_table(id, mode));
        const int pick = std::min<int>(std::max(0, env_int("ACLFFT_MFMA_VARIANT", 0)), (int)ids.size() - 1);
        s.kernel = ids[pick];
        s.tw_table = s.candidate_tables[pick];
        s.aux_table = roots_table(n, n);
        s.src = src;
        s.dst = dst;
        s.args.batch = batch;
        s.args.idist = n;
        s.args.odist = n;
        s.args.in_sign = s.args.out_sign = s.args.scale = 1.f;
        s.conj_in = s.conj_out = conj;
        e->steps_.push_back(s);
        return ACLFFT_SUCCESS;
    }

    // 2D C2C in an MFMA precision: images of nx rows and ny columns with nx, ny in {16, 32, 64}, each
    // as the two dense DFT stages of an MFMA kernel with r1 = ny, r2 = nx, no inter-stage twiddle
    // (a table of ones) and the output stored in the image layout (OFX_FLAG_MFMA_2D): the column
    // DFTs are F_nx X, the row DFTs (F_nx X) F_ny, the reference's P X Q
    // (src/fft2_d/arch32/dd/fft2_d_dd.cpp:31-66) on the matrix cores.
    static std::vector<int> mfma_2d_kernels_for(long long nx, long long ny) {
        std::vector<int> ids, tiled;
        for (int i = 1; i < ofx_mfma_kernel_count(); ++i) {
            ofx_mfma_info mi;
            if (ofx_mfma_kernel_info_get(i, &mi) == 0 && mi.r1 == ny && mi.r2 == nx) {
                ids.push_back(i);
                tiled.push_back(mi.tiled);
            }
        }
        if (nx * ny > 2048) {  // 64 x 64: the tiled form first, as for the 1D 4096-point transform
            std::vector<int> sorted;
            for (int want = 1; want >= 0; --want)
                for (size_t k = 0; k < ids.size(); ++k)
                    if (tiled[k] == want) sorted.push_back(ids[k]);
            ids = sorted;
        }
        return ids;
    }
    aclfftResult c2c_2d_mfma(long long nx, long long ny, long long images, int mode) {
        const std::vector<int> ids = mfma_2d_kernels_for(nx, ny);
        if (ids.empty() || ofx_mfma_kernel_const_bytes(ids[0], mode) <= 0) return ACLFFT_NOT_SUPPORTED;
        const long long n = nx * ny;
        Step s;
        s.type = Step::MFMA;
        s.n = (int)n;
        s.mfma_mode = mode;
        s.candidates = ids;
        for (int id : ids) s.candidate_tables.push_back(mfma_table(id, mode));
        const int pick = std::min<int>(std::max(0, env_int("ACLFFT_MFMA_VARIANT", 0)), (int)ids.size() - 1);
        s.kernel = ids[pick];
        s.tw_table = s.candidate_tables[pick];
        s.aux_table = roots_table(1, n);  // W_1^k = 1: no inter-stage twiddle
        s.src = BUF_IN;
        s.dst = BUF_OUT;
        s.args.batch = images;
        s.args.idist = n;
        s.args.odist = n;
        s.args.flags = OFX_FLAG_MFMA_2D;
        s.args.in_sign = s.args.out_sign = s.args.scale = 1.f;
        s.conj_in = s.conj_out = true;
        e->steps_.push_back(s);
        return ACLFFT_SUCCESS;
    }

    // Real transforms in an MFMA precision (even n, n/2 with an MFMA kernel): R2C as the n/2-point
    // MFMA transform of the real samples viewed as complex, into the workspace, then the Hermitian
    // unpack; C2R as the pre-pack into the workspace (conjugated), then the forward MFMA transform
    // with a conjugated output.
    aclfftResult r2c_mfma(long long n, long long batch, int mode) {
        const long long m = n / 2;
        if (n % 2 || mfma_kernels_for(m).empty()) return ACLFFT_NOT_SUPPORTED;
        e->in_place_unsafe_ = true;
        aclfftResult r = c2c_mfma(m, batch, mode, BUF_IN, BUF_WORK1, false);
        if (r != ACLFFT_SUCCESS) return r;
        Step u;
        u.type = Step::R2C_UNPACK;
        u.src = BUF_WORK1;
        u.dst = BUF_OUT;
        u.m = m;
        u.args.batch = batch;
        u.args.scale = 1.f;
        u.aux_table = roots_table(n, n / 2 + 1);
        e->steps_.push_back(u);
        e->work1_bytes_ = std::max(e->work1_bytes_, (size_t)(m * batch * 8));
        return ACLFFT_SUCCESS;
    }
    aclfftResult c2r_mfma(long long n, long long batch, int mode) {
        const long long m = n / 2;
        if (n % 2 || mfma_kernels_for(m).empty()) return ACLFFT_NOT_SUPPORTED;
        e->in_place_unsafe_ = true;
        Step pk;
        pk.type = Step::C2R_PACK;
        pk.src = BUF_IN;
        pk.dst = BUF_WORK1;
        pk.m = m;
        pk.args.batch = batch;
        pk.aux_table = roots_table(n, n / 2 + 1);
        e->steps_.push_back(pk);
        e->work1_bytes_ = std::max(e->work1_bytes_, (size_t)(m * batch * 8));
        aclfftResult r = c2c_mfma(m, batch, mode, BUF_WORK1, BUF_OUT, false);
        if (r != ACLFFT_SUCCESS) return r;
        e->steps_.back().args.out_sign = -1.f;
        return ACLFFT_SUCCESS;
    }

    // Uneven batch, one launch per length: every length with a compiled tile-list kernel gets one
    // step over the tiles (<= TPB contiguous transforms of one group) of all its groups; any other
    // length runs as ordinary steps per group.  Each length runs with its own kernel's register
    // and LDS budget, which a single persistent kernel has to size for its largest class.
    aclfftResult uneven_grouped(const std::vector<int>& len, const std::vector<int>& bat) {
        std::vector<long long> offs(len.size());
        long long off = 0;
        for (size_t g = 0; g < len.size(); ++g) {
            offs[g] = off;
            off += (long long)len[g] * bat[g];
        }
        e->in_bytes_ = e->out_bytes_ = (size_t)off * 8;
        aclfftResult r = per_length_steps(len, bat, offs, nullptr, true);
        if (r != ACLFFT_SUCCESS) return r;
        // steps of lengths without a tile kernel (lane 0) first, so that the lanes fork once
        std::stable_partition(e->steps_.begin(), e->steps_.end(), [](const Step& a) { return a.lane == 0; });
        return ACLFFT_SUCCESS;
    }

    // One tile launch per distinct length over its groups (groups g with take[g] set, or all),
    // largest total work first; lengths without a tile kernel run one step per group.  With
    // `lanes`, ACLFFT_UNEVEN_LANES spreads the launches over side streams.
    aclfftResult per_length_steps(const std::vector<int>& len, const std::vector<int>& bat,
                                  const std::vector<long long>& offs, const std::vector<char>* take, bool lanes_ok) {
        // distinct lengths, largest total work first
        std::map<int, double> work;
        for (size_t g = 0; g < len.size(); ++g)
            if (take == nullptr || (*take)[g]) work[len[g]] += (double)len[g] * bat[g];
        std::vector<int> lens;
        for (auto& kv : work) lens.push_back(kv.first);
        std::stable_sort(lens.begin(), lens.end(), [&](int a, int b) { return work[a] > work[b]; });
        std::vector<double> lane_load;
        for (int n : lens) {
            auto ids = kernels_for(OFX_KIND_C2C_TILES, n, true);
            if (ids.empty()) {
                for (size_t g = 0; g < len.size(); ++g) {
                    if (len[g] != n || (take != nullptr && !(*take)[g])) continue;
                    const size_t first = e->steps_.size();
                    aclfftResult r = c2c_rows(n, bat[g], BUF_IN, BUF_OUT, true, true);
                    if (r != ACLFFT_SUCCESS) return r;
                    e->steps_.back().final_out = true;
                    for (size_t k = first; k < e->steps_.size(); ++k) {
                        if (e->steps_[k].src == BUF_IN || e->steps_[k].src == BUF_OUT)
                            e->steps_[k].src_off = (size_t)offs[g] * 8;
                        if (e->steps_[k].dst == BUF_OUT) e->steps_[k].dst_off = (size_t)offs[g] * 8;
                    }
                }
                continue;
            }
            Step s = kernel_step(OFX_KIND_C2C_TILES, n, ids);
            // the tile list is cut for one TPB: only candidates with that TPB stay tunable
            const int tpb = registry()[s.kernel].transforms_per_block;
            std::vector<int> cand, tabs;
            for (size_t c = 0; c < s.candidates.size(); ++c)
                if (registry()[s.candidates[c]].transforms_per_block == tpb) {
                    cand.push_back(s.candidates[c]);
                    tabs.push_back(s.candidate_tables[c]);
                }
            s.candidates = cand;
            s.candidate_tables = tabs;
            std::vector<ofx_batch_tile> tiles;
            for (size_t g = 0; g < len.size(); ++g) {
                if (len[g] != n || (take != nullptr && !(*take)[g])) continue;
                for (long long i = 0; i < bat[g]; i += tpb) {
                    ofx_batch_tile t;
                    t.cls = 0;
                    t.count = (int)std::min<long long>(tpb, bat[g] - i);
                    t.offset = offs[g] + i * n;
                    tiles.push_back(t);
                }
            }
            TableSpec tt;
            tt.kind = TableSpec::BLOB;
            tt.bytes = tiles.size() * sizeof(ofx_batch_tile);
            tt.blob.resize(tt.bytes);
            std::memcpy(tt.blob.data(), tiles.data(), tt.bytes);
            s.tiles_table = e->add_table(tt);
            s.args.batch = (long long)tiles.size();
            s.conj_in = s.conj_out = true;
            s.final_out = true;
            // longest-processing-time assignment of lengths to side-stream lanes, off by default:
            // the fork / join of the side streams costs more than the overlap gains (2 to 8 lanes
            // measured 1.2-3x slower than the plan's own stream for one to three lengths)
            const int lanes = lanes_ok ? std::max(0, std::min(8, env_int("ACLFFT_UNEVEN_LANES", 0))) : 0;
            if (lanes > 1) {
                lane_load.resize(lanes, 0.0);
                const size_t k = std::min_element(lane_load.begin(), lane_load.end()) - lane_load.begin();
                lane_load[k] += work[n];
                s.lane = (int)k + 1;
            }
            e->steps_.push_back(s);
        }
        return ACLFFT_SUCCESS;
    }

    // Groups of (length, batch) stored one after the other.  ACLFFT_UNEVEN_MODE=1 selects the
    // persistent kernel (below), 2 the per-length launches (uneven_grouped).  By default the
    // per-length launches are used for a single length, or when every length carries at least
    // 2^24 points (each launch then amortises its ramp-up and tail, and runs with its own
    // kernel's register and LDS budget), the persistent kernel otherwise (several lengths with
    // less work each, whose tiles share one launch).  Measured: one length of 5-134 M points,
    // per-length 3-11% faster; 1024 + 4096 with 17 and 67 M points, 8% faster; 2048 + 4096 + 6561
    // with 5-11 M points each, persistent 13% faster; 512 groups of 8 lengths, 29% faster.
    aclfftResult uneven(const std::vector<int>& len, const std::vector<int>& bat) {
        int mode = e->desc_.uneven_form ? e->desc_.uneven_form : env_int("ACLFFT_UNEVEN_MODE", 0);
        if (mode != 1 && mode != 2) {
            std::map<int, long long> pts;
            for (size_t g = 0; g < len.size(); ++g) pts[len[g]] += (long long)len[g] * bat[g];
            long long least = -1;
            for (auto& kv : pts) least = least < 0 ? kv.second : std::min(least, kv.second);
            mode = (pts.size() == 1 || least >= (1LL << 24)) ? 2 : 1;
        }
        return mode == 2 ? uneven_grouped(len, bat) : uneven_persistent(len, bat);
    }

    // Lengths with a compiled batch class become tiles of the persistent kernel (one launch); any
    // other length runs as its own step.
    // Uneven plan in an MFMA precision: every length with a tile-capable MFMA kernel (registry
    // entries other than 0) gets one MFMA launch over its groups' tiles (up to 16 contiguous
    // transforms each, one wave per transform); the other lengths run as in the FP32 plan (the
    // persistent kernel and, for lengths without a class, their own steps).
    aclfftResult uneven_mfma(const std::vector<int>& len, const std::vector<int>& bat, int mode) {
        constexpr int kTile = 16;
        std::vector<long long> offs(len.size());
        long long off = 0;
        std::vector<char> on_mfma(len.size(), 0);
        std::map<int, std::vector<int>> ids_of;
        for (size_t g = 0; g < len.size(); ++g) {
            offs[g] = off;
            off += (long long)len[g] * bat[g];
            if (ids_of.count(len[g]) == 0) {
                ids_of[len[g]] = mfma_tile_kernels_for(len[g]);
            }
            on_mfma[g] = ids_of[len[g]].empty() ? 0 : 1;
        }
        bool any_fp32 = false;
        for (char c : on_mfma) any_fp32 = any_fp32 || !c;
        if (any_fp32) {
            aclfftResult r = uneven_persistent(len, bat, &on_mfma);
            if (r != ACLFFT_SUCCESS) return r;
        }
        e->in_bytes_ = e->out_bytes_ = (size_t)off * 8;
        // one step per MFMA length, largest total work first
        std::map<int, double> work;
        for (size_t g = 0; g < len.size(); ++g)
            if (on_mfma[g]) work[len[g]] += (double)len[g] * bat[g];
        std::vector<int> lens;
        for (auto& kv : work) lens.push_back(kv.first);
        std::stable_sort(lens.begin(), lens.end(), [&](int a, int b) { return work[a] > work[b]; });
        for (int n : lens) {
            std::vector<ofx_batch_tile> tiles;
            for (size_t g = 0; g < len.size(); ++g) {
                if (!on_mfma[g] || len[g] != n) continue;
                for (long long i = 0; i < bat[g]; i += kTile) {
                    ofx_batch_tile t;
                    t.cls = 0;
                    t.count = (int)std::min<long long>(kTile, bat[g] - i);
                    t.offset = offs[g] + i * n;
                    tiles.push_back(t);
                }
            }
            Step s;
            s.type = Step::MFMA;
            s.n = n;
            s.mfma_mode = mode;
            s.candidates = ids_of[n];
            for (int id : s.candidates) s.candidate_tables.push_back(mfma_table(id, mode));
            const int pick = std::min<int>(std::max(0, env_int("ACLFFT_MFMA_VARIANT", 0)), (int)s.candidates.size() - 1);
            s.kernel = s.candidates[pick];
            s.tw_table = s.candidate_tables[pick];
            s.aux_table = roots_table(n, n);
            TableSpec tt;
            tt.kind = TableSpec::BLOB;
            tt.bytes = tiles.size() * sizeof(ofx_batch_tile);
            tt.blob.resize(tt.bytes);
            std::memcpy(tt.blob.data(), tiles.data(), tt.bytes);
            s.tiles_table = e->add_table(tt);
            s.args.batch = (long long)tiles.size();
            s.args.idist = s.args.odist = n;
            s.args.in_sign = s.args.out_sign = s.args.scale = 1.f;
            s.conj_in = s.conj_out = true;
            s.final_out = true;
            e->steps_.push_back(s);
        }
        return ACLFFT_SUCCESS;
    }

    aclfftResult uneven_persistent(const std::vector<int>& len, const std::vector<int>& bat,
                                   const std::vector<char>* skip = nullptr) {
        struct ClassInfo {
            ofx_batch_class_info info;
            bool used = false;
        };
        const int cset = e->desc_.class_set >= 0
                             ? e->desc_.class_set
                             : std::min(std::max(0, env_int("ACLFFT_PERSISTENT_CLASS_SET", 0)),
                                        ofx_batch_class_set_count() - 1);
        std::vector<ClassInfo> cls(ofx_batch_class_count());
        for (int c = 0; c < (int)cls.size(); ++c)
            if (ofx_batch_class_info_get_set(cset, c, &cls[c].info) != 0) return ACLFFT_INVALID_VALUE;
        std::vector<ofx_batch_tile> tiles;
        std::vector<double> cost;
        long long off = 0;
        size_t total = 0;
        std::vector<long long> offs(len.size());
        std::vector<char> own(len.size(), 0);  // groups of lengths without a class in this set
        for (size_t g = 0; g < len.size(); ++g) {
            const long long n = len[g], b = bat[g];
            offs[g] = off;
            if (skip != nullptr && (*skip)[g]) {  // planned elsewhere; only its offset counts here
                off += n * b;
                total += (size_t)(n * b);
                continue;
            }
            int c = -1;
            for (int k = 0; k < (int)cls.size(); ++k)
                if (cls[k].info.n == n) c = k;
            if (c >= 0) {
                cls[c].used = true;
                const int tpt = cls[c].info.transforms_per_tile;
                for (long long i = 0; i < b; i += tpt) {
                    ofx_batch_tile t;
                    t.cls = c;
                    t.count = (int)std::min<long long>(tpt, b - i);
                    t.offset = off + i * n;
                    tiles.push_back(t);
                    cost.push_back((double)t.count * n * std::log2((double)n + 1.0));
                }
            } else {
                own[g] = 1;
            }
            off += n * b;
            total += (size_t)(n * b);
        }
        e->in_bytes_ = e->out_bytes_ = total * 8;
        {
            aclfftResult r = per_length_steps(len, bat, offs, &own, false);
            if (r != ACLFFT_SUCCESS) return r;
        }
        if (tiles.empty()) return ACLFFT_SUCCESS;
        // longest-processing-time order
        std::vector<size_t> order(tiles.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
        Step s;
        s.type = Step::PERSISTENT;
        s.conj_in = s.conj_out = true;
        s.final_out = true;
        s.b.scale = 1.f;
        s.b.class_set = cset;
        s.b.schedule = e->desc_.sched >= 0 ? (e->desc_.sched ? 1 : 0)
                                           : (env_int("ACLFFT_PERSISTENT_SCHEDULE", 1) == 0 ? 0 : 1);
        s.blocks_per_cu = e->desc_.blocks_per_cu > 0 ? e->desc_.blocks_per_cu
                                                     : std::max(0, env_int("ACLFFT_PERSISTENT_BLOCKS_PER_CU", 0));
        // small-tile prologue: the walk's first 3 tiles per workgroup are the smallest ones, so the
        // workgroups drift apart before the long tiles (512-group batch: 83.0-83.8 vs 90.3 us;
        // 1 per workgroup 88.5, 2 88.9-89.2, 4 84.7 us, the whole list reversed 87.0-87.3 us)
        {
            const int env_pro = env_int("ACLFFT_PERSISTENT_PROLOGUE", -1);
            s.b.prologue_rounds = e->desc_.prologue >= 0 ? e->desc_.prologue : env_pro >= 0 ? env_pro : 3;
        }
        {
            const int env_stream = env_int("ACLFFT_PERSISTENT_STREAM", -1);
            s.b.stream = e->desc_.stream >= 0 ? (e->desc_.stream ? 1 : 0)
                         : env_stream >= 0    ? (env_stream ? 1 : 0)
                                              : (total * 16 >= kStreamBytes ? 1 : 0);
        }
        for (int c = 0; c < (int)cls.size(); ++c) {
            s.class_tables[c] = -1;
            if (cls[c].used) s.class_tables[c] = pass_table_for(cls[c].info.npass, cls[c].info.radices, cls[c].info.tw_len);
        }
        TableSpec tt;
        tt.kind = TableSpec::BLOB;
        tt.bytes = tiles.size() * sizeof(ofx_batch_tile);
        tt.blob.resize(tt.bytes);
        for (size_t i = 0; i < order.size(); ++i)
            std::memcpy(tt.blob.data() + i * sizeof(ofx_batch_tile), &tiles[order[i]], sizeof(ofx_batch_tile));
        s.tiles_table = e->add_table(tt);
        TableSpec ct;
        ct.kind = TableSpec::BLOB;
        ct.bytes = 4;
        ct.blob.assign(4, 0);
        s.counter_table = e->add_table(ct);
        s.b.ntiles = (int)tiles.size();
        e->steps_.insert(e->steps_.begin(), s);
        return ACLFFT_SUCCESS;
    }

    // complex128 rows: a compiled kernel where the registry has one, else the four-step over
    // compiled complex128 column and row-step kernels; other lengths plan successfully and report
    // NOT_IMPLEMENTED at execution.
    aclfftResult z2z_rows(long long n, long long batch) {
        return c2c_rows(n, batch, BUF_IN, BUF_OUT, true, true, true);
    }
    aclfftResult d2z(long long n, long long batch) {
        e->in_place_unsafe_ = true;
        const bool even = (n % 2) == 0;
        auto ids = kernels_for(even ? OFX_KIND_D2Z : OFX_KIND_D2Z_ODD, even ? n / 2 : n, true);
        if (ids.empty() && even && n / 2 > OFX_GENERIC_MAX_ELEMS_Z) {
            // large even n: the complex128 four-step of the n/2 packed samples into the workspace,
            // then the Hermitian unpack
            const long long m = n / 2;
            aclfftResult r = c2c_rows(m, batch, BUF_IN, BUF_WORK1, false, false, true);
            if (r != ACLFFT_SUCCESS) return r;
            Step u;
            u.type = Step::R2C_UNPACK;
            u.kind = OFX_KIND_D2Z;
            u.src = BUF_WORK1;
            u.dst = BUF_OUT;
            u.m = m;
            u.args.batch = batch;
            u.args.scale = 1.f;
            u.args.scale64 = 1.0;
            u.aux_table = roots_table_z(n, n / 2 + 1);
            e->steps_.push_back(u);
            e->work1_bytes_ = std::max(e->work1_bytes_, (size_t)(m * batch) * 16);
            return ACLFFT_SUCCESS;
        }
        if (ids.empty()) {
            // runtime-radix kernel: odd n as a full complex transform of the real samples (half
            // spectrum stored), even n as the n/2-point transform of the packed samples with the
            // Hermitian unpack fused into its store
            Step g;
            if (!generic_step(even ? n / 2 : n, &g, true)) return ACLFFT_NOT_IMPLEMENTED;
            g.args.batch = batch;
            g.args.istride = 1;
            g.args.ostride = 1;
            g.args.odist = n / 2 + 1;
            if (even) {
                g.g.in_mode = OFX_IN_COMPLEX;
                g.g.out_mode = OFX_OUT_R2C_UNPACK;
                g.args.idist = n / 2;
                g.aux_table = roots_table_z(n, n / 2 + 1);
            } else {
                g.g.in_mode = OFX_IN_REAL;
                g.g.out_mode = OFX_OUT_HALF;
                g.args.idist = n;
            }
            e->steps_.push_back(g);
            return ACLFFT_SUCCESS;
        }
        Step s = kernel_step(even ? OFX_KIND_D2Z : OFX_KIND_D2Z_ODD, even ? n / 2 : n, ids);
        s.args.batch = batch;
        s.args.idist = n;
        s.args.odist = n / 2 + 1;
        if (even) s.aux_table = roots_table_z(n, n / 2 + 1);
        e->steps_.push_back(s);
        return ACLFFT_SUCCESS;
    }
    aclfftResult z2d(long long n, long long batch) {
        e->in_place_unsafe_ = true;
        const bool even = (n % 2) == 0;
        auto ids = kernels_for(even ? OFX_KIND_Z2D : OFX_KIND_Z2D_ODD, even ? n / 2 : n, true);
        if (ids.empty() && even && n / 2 > OFX_GENERIC_MAX_ELEMS_Z) {
            // large even n: Hermitian pre-pack into the workspace (conjugated), then the complex128
            // four-step forward with a conjugated output
            const long long m = n / 2;
            Step pk;
            pk.type = Step::C2R_PACK;
            pk.kind = OFX_KIND_Z2D;
            pk.src = BUF_IN;
            pk.dst = BUF_WORK1;
            pk.m = m;
            pk.args.batch = batch;
            pk.aux_table = roots_table_z(n, n / 2 + 1);
            e->steps_.push_back(pk);
            e->work1_bytes_ = std::max(e->work1_bytes_, (size_t)(m * batch) * 16);
            aclfftResult r = c2c_rows(m, batch, BUF_WORK1, BUF_OUT, false, false, true);
            if (r != ACLFFT_SUCCESS) return r;
            e->steps_.back().args.out_sign = -1.f;
            return ACLFFT_SUCCESS;
        }
        if (ids.empty()) {
            // runtime-radix kernel: odd n from the Hermitian expansion of the half spectrum, even n
            // as the pre-pack fused into the load of an n/2-point forward transform with a
            // conjugated store
            Step g;
            if (!generic_step(even ? n / 2 : n, &g, true)) return ACLFFT_NOT_IMPLEMENTED;
            g.args.batch = batch;
            g.args.idist = n / 2 + 1;
            g.args.istride = 1;
            g.args.ostride = 1;
            if (even) {
                g.g.in_mode = OFX_IN_C2R_PACK;
                g.g.out_mode = OFX_OUT_COMPLEX;
                g.args.out_sign = -1.f;
                g.args.odist = n / 2;
                g.aux_table = roots_table_z(n, n / 2 + 1);
            } else {
                g.g.in_mode = OFX_IN_HERMITIAN;
                g.g.out_mode = OFX_OUT_REAL;
                g.args.odist = n;
            }
            e->steps_.push_back(g);
            return ACLFFT_SUCCESS;
        }
        Step s = kernel_step(even ? OFX_KIND_Z2D : OFX_KIND_Z2D_ODD, even ? n / 2 : n, ids);
        s.args.batch = batch;
        s.args.idist = n / 2 + 1;
        s.args.odist = n;
        if (even) s.aux_table = roots_table_z(n, n / 2 + 1);
        e->steps_.push_back(s);
        return ACLFFT_SUCCESS;
    }

    aclfftResult r2c(long long n, long long batch) {
        const long long h = n / 2 + 1;
        e->in_place_unsafe_ = true;
        const bool even = (n % 2) == 0;
        if (n == 1) {
            // X[0] = x[0]: the odd-n path with the length-1 generic step (no passes)
        }
        auto ids = kernels_for(even ? OFX_KIND_R2C : OFX_KIND_R2C_ODD, even ? n / 2 : n, true);
        if (!ids.empty()) {
            Step s = kernel_step(even ? OFX_KIND_R2C : OFX_KIND_R2C_ODD, even ? n / 2 : n, ids);
            s.args.batch = batch;
            s.args.idist = n;
            s.args.odist = h;
            if (even) s.aux_table = roots_table(n, n / 2 + 1);
            e->steps_.push_back(s);
            return ACLFFT_SUCCESS;
        }
        const long long m = even ? n / 2 : n;
        Step s;
        if (generic_step(m, &s)) {
            s.args.batch = batch;
            s.args.istride = 1;
            s.args.ostride = 1;
            if (even) {
                s.g.in_mode = OFX_IN_COMPLEX;
                s.g.out_mode = OFX_OUT_R2C_UNPACK;
                s.args.idist = m;  // complex units over the real data
                s.args.odist = h;
                s.aux_table = roots_table(n, n / 2 + 1);
            } else {
                s.g.in_mode = OFX_IN_REAL;
                s.g.out_mode = OFX_OUT_HALF;
                s.args.idist = n;
                s.args.odist = h;
            }
            e->steps_.push_back(s);
            return ACLFFT_SUCCESS;
        }
        if (!even) return ACLFFT_NOT_IMPLEMENTED;
        // large even n: M-point four-step into the workspace, then the unpack pass
        const size_t before = e->steps_.size();
        aclfftResult r = c2c_rows(m, batch, BUF_IN, BUF_WORK1, false, false);
        if (r != ACLFFT_SUCCESS) return r;
        (void)before;
        Step u;
        u.type = Step::R2C_UNPACK;
        u.src = BUF_WORK1;
        u.dst = BUF_OUT;
        u.m = m;
        u.args.batch = batch;
        u.args.scale = 1.f;
        u.aux_table = roots_table(n, n / 2 + 1);
        e->steps_.push_back(u);
        e->work1_bytes_ = std::max(e->work1_bytes_, (size_t)(m * batch * 8));
        return ACLFFT_SUCCESS;
    }

    aclfftResult c2r(long long n, long long batch) {
        const long long h = n / 2 + 1;
        e->in_place_unsafe_ = true;
        const bool even = (n % 2) == 0;
        auto ids = kernels_for(even ? OFX_KIND_C2R : OFX_KIND_C2R_ODD, even ? n / 2 : n, true);
        if (!ids.empty()) {
            Step s = kernel_step(even ? OFX_KIND_C2R : OFX_KIND_C2R_ODD, even ? n / 2 : n, ids);
            s.args.batch = batch;
            s.args.idist = h;
            s.args.odist = n;
            if (even) s.aux_table = roots_table(n, n / 2 + 1);
            e->steps_.push_back(s);
            return ACLFFT_SUCCESS;
        }
        const long long m = even ? n / 2 : n;
        Step s;
        if (generic_step(m, &s)) {
            s.args.batch = batch;
            s.args.idist = h;
            s.args.istride = 1;
            s.args.ostride = 1;
            if (even) {
                s.g.in_mode = OFX_IN_C2R_PACK;
                s.g.out_mode = OFX_OUT_COMPLEX;
                s.args.out_sign = -1.f;
                s.args.odist = m;  // complex units over the real output
                s.aux_table = roots_table(n, n / 2 + 1);
            } else {
                s.g.in_mode = OFX_IN_HERMITIAN;
                s.g.out_mode = OFX_OUT_REAL;
                s.args.odist = n;
            }
            e->steps_.push_back(s);
            return ACLFFT_SUCCESS;
        }
        if (!even) return ACLFFT_NOT_IMPLEMENTED;
        Step pk;
        pk.type = Step::C2R_PACK;
        pk.src = BUF_IN;
        pk.dst = BUF_WORK1;
        pk.m = m;
        pk.args.batch = batch;
        pk.aux_table = roots_table(n, n / 2 + 1);
        e->steps_.push_back(pk);
        e->work1_bytes_ = std::max(e->work1_bytes_, (size_t)(m * batch * 8));
        // the packed data is already conjugated: forward transform, conjugated output
        const size_t first = e->steps_.size();
        aclfftResult r = c2c_rows(m, batch, BUF_WORK1, BUF_OUT, false, false);
        if (r != ACLFFT_SUCCESS) return r;
        e->steps_.back().args.out_sign = -1.f;
        (void)first;
        return ACLFFT_SUCCESS;
    }
};

int Executor::add_table(const TableSpec& t) {
    TableSpec c = t;
    c.offset = table_bytes_;
    table_bytes_ += align256(c.bytes);
    tables_.push_back(c);
    return (int)tables_.size() - 1;
}

aclfftResult Executor::build(const PlanDesc& d, Executor** out) {
    *out = nullptr;
    Executor* e = new (std::nothrow) Executor();
    if (e == nullptr) return ACLFFT_ALLOC_FAILED;
    e->desc_ = d;
    Builder b{e};
    aclfftResult r = ACLFFT_NOT_IMPLEMENTED;
    if (d.precision != 0 && !(d.rank == 1 && !d.vertical &&
                              (d.type == ACLFFT_C2C || d.type == ACLFFT_R2C || d.type == ACLFFT_C2R)) &&
        !(d.rank == 2 && d.type == ACLFFT_C2C)) {
        delete e;
        return ACLFFT_NOT_SUPPORTED;
    }
    if (d.uneven) {
        if (d.type != ACLFFT_C2C) {
            r = ACLFFT_NOT_SUPPORTED;
        } else if (d.precision != 0) {
            // an MFMA precision needs at least one length with a tile-list MFMA kernel
            bool any = false;
            for (int n : d.ulen) any = any || !Builder::mfma_tile_kernels_for(n).empty();
            r = any ? b.uneven_mfma(d.ulen, d.ubatch, d.precision) : ACLFFT_NOT_SUPPORTED;
        } else {
            r = b.uneven(d.ulen, d.ubatch);
        }
        // replaying a graph costs a few microseconds per launch on the device: worth it only when
        // it replaces several launches (e.g. the ordinary steps of lengths without a class)
        if (r == ACLFFT_SUCCESS) e->use_graphs_ = e->steps_.size() > 3;
    } else if (d.rank == 1 && !d.vertical) {
        const long long n = d.n0;
        switch (d.type) {
            case ACLFFT_C2C:
                e->in_bytes_ = e->out_bytes_ = (size_t)n * d.batch * 8;
                r = d.precision ? b.c2c_mfma(n, d.batch, d.precision) : b.c2c_rows(n, d.batch, BUF_IN, BUF_OUT, true, true);
                break;
            case ACLFFT_R2C:
                e->in_bytes_ = (size_t)n * d.batch * 4;
                e->out_bytes_ = (size_t)(n / 2 + 1) * d.batch * 8;
                r = d.precision ? b.r2c_mfma(n, d.batch, d.precision) : b.r2c(n, d.batch);
                break;
            case ACLFFT_C2R:
                e->in_bytes_ = (size_t)(n / 2 + 1) * d.batch * 8;
                e->out_bytes_ = (size_t)n * d.batch * 4;
                r = d.precision ? b.c2r_mfma(n, d.batch, d.precision) : b.c2r(n, d.batch);
                break;
            case ACLFFT_Z2Z:
                e->in_bytes_ = e->out_bytes_ = (size_t)n * d.batch * 16;
                r = b.z2z_rows(n, d.batch);
                break;
            case ACLFFT_D2Z:
                e->in_bytes_ = (size_t)n * d.batch * 8;
                e->out_bytes_ = (size_t)(n / 2 + 1) * d.batch * 16;
                r = b.d2z(n, d.batch);
                break;
            case ACLFFT_Z2D:
                e->in_bytes_ = (size_t)(n / 2 + 1) * d.batch * 16;
                e->out_bytes_ = (size_t)n * d.batch * 8;
                r = b.z2d(n, d.batch);
                break;
            default:
                r = ACLFFT_NOT_SUPPORTED;
        }
        if (r == ACLFFT_SUCCESS && d.precision == 0) b.stream_stores(e->out_bytes_);
    } else if (d.rank == 1 && d.vertical && (d.type == ACLFFT_C2C || d.type == ACLFFT_Z2Z)) {
        // [n][cols] row-major, one transform per column (reference layout: VERTICAL, golden
        // src/fft1_d/tests/fft1_d_data/gen_data.py:79-81)
        const bool z = d.type == ACLFFT_Z2Z;
        const long long n = d.n0, cols = d.batch;
        e->in_bytes_ = e->out_bytes_ = (size_t)n * cols * (z ? 16 : 8);
        r = b.c2c_columns(n, cols, 1, BUF_IN, BUF_OUT, true, true, z);
    } else if ((d.rank == 2 || d.rank == 3) && (d.type == ACLFFT_C2C || d.type == ACLFFT_Z2Z)) {
        const bool z = d.type == ACLFFT_Z2Z;
        // axes slowest first; length-1 axes are dropped (a 64x1x32 volume is a 64x32 image)
        std::vector<long long> dims;
        long long total = 1;
        const int lens[3] = {d.n0, d.n1, d.n2};
        for (int i = 0; i < d.rank; ++i) {
            total *= lens[i];
            if (lens[i] > 1) dims.push_back(lens[i]);
        }
        e->in_bytes_ = e->out_bytes_ = (size_t)total * d.batch * (z ? 16 : 8);
        if (dims.empty()) dims.push_back(1);
        // rows along the fastest axis (contiguous), then every slower axis as columns in place:
        // axis i has length dims[i], stride inner = product of the faster axes, and
        // outer = batch * product of the slower axes images
        const int k = (int)dims.size();
        long long inner = dims[k - 1];
        if (d.precision != 0) {
            // MFMA precision: 2D C2C images of 16, 32 or 64 points per axis
            r = (k == 2 && !z) ? b.c2c_2d_mfma(dims[0], dims[1], total / (dims[0] * dims[1]) * d.batch, d.precision)
                               : ACLFFT_NOT_SUPPORTED;
        } else if (k == 2 && !z && b.fused_2d(dims[0], dims[1], total / (dims[0] * dims[1]) * d.batch)) {
            // images that fit in LDS: rows and columns in one pass over HBM
            r = ACLFFT_SUCCESS;
        } else {
        r = b.c2c_rows(dims[k - 1], total / dims[k - 1] * d.batch, BUF_IN, BUF_OUT, true, k == 1, z);
        for (int i = k - 2; i >= 0 && r == ACLFFT_SUCCESS; --i) {
            const long long outer = total / (inner * dims[i]) * d.batch;
            r = b.c2c_columns(dims[i], inner, outer, BUF_OUT, BUF_OUT, false, i == 0, z);
            inner *= dims[i];
        }
        }
        if (r == ACLFFT_SUCCESS && !z) b.stream_final_multi(e->in_bytes_, e->out_bytes_);
    }
    if (r != ACLFFT_SUCCESS) {
        delete e;
        return r;
    }
    const aclfftType t = d.type;
    if (!d.hermitian_weight && (t == ACLFFT_R2C || t == ACLFFT_C2R || t == ACLFFT_D2Z || t == ACLFFT_Z2D)) {
        // the adjoint execution (the opposite real type) runs on the same caller work area
        PlanDesc a = d;
        a.type = t == ACLFFT_R2C ? ACLFFT_C2R : t == ACLFFT_C2R ? ACLFFT_R2C : t == ACLFFT_D2Z ? ACLFFT_Z2D : ACLFFT_D2Z;
        a.hermitian_weight = true;
        Executor* twin = nullptr;
        if (build(a, &twin) == ACLFFT_SUCCESS) {
            e->twin_work_bytes_ = twin->work_bytes();
            delete twin;
        }
    }
    if (d.hermitian_weight)
        for (Step& st : e->steps_)
            if (st.type != Step::COPY) st.args.flags |= OFX_FLAG_HERMITIAN_WEIGHT;
    if (d.scale != 1.f && d.uneven) {
        // every group's output region is written last by its own step (a tile step of its
        // length, the persistent launch, or the last step of the group's own launch sequence)
        for (Step& st : e->steps_) {
            if (!st.final_out) continue;
            if (st.type == Step::PERSISTENT) {
                st.b.scale *= d.scale;
            } else {
                st.args.scale *= d.scale;
                st.args.scale64 = (double)st.args.scale;
            }
        }
    } else if (d.scale != 1.f) {
        // fold the normalisation into the last launch that writes the output
        for (auto it = e->steps_.rbegin(); it != e->steps_.rend(); ++it)
            if (it->dst == BUF_OUT && it->type != Step::COPY && it->type != Step::PERSISTENT) {
                it->args.scale *= d.scale;
                it->args.scale64 = (double)it->args.scale;
                break;
            }
    }
    *out = e;
    return ACLFFT_SUCCESS;
}

Executor* Executor::adjoint() {
    const aclfftType t = desc_.type;
    if (adjoint_ == nullptr && (t == ACLFFT_R2C || t == ACLFFT_C2R || t == ACLFFT_D2Z || t == ACLFFT_Z2D)) {
        PlanDesc d = desc_;
        d.type = t == ACLFFT_R2C ? ACLFFT_C2R : t == ACLFFT_C2R ? ACLFFT_R2C : t == ACLFFT_D2Z ? ACLFFT_Z2D : ACLFFT_D2Z;
        d.hermitian_weight = true;
        Executor* e = nullptr;
        if (build(d, &e) == ACLFFT_SUCCESS) {
            adjoint_ = e;
            // a caller-managed workspace is shared with the adjoint (same stream, never concurrent)
            e->auto_alloc_ = auto_alloc_;
            if (user_work_ != nullptr && e->work_bytes() <= work_bytes()) (void)e->set_work_area(user_work_);
        }
    }
    return adjoint_;
}

Executor::~Executor() {
    int cur = -1;
    if (device_ >= 0 && hipGetDevice(&cur) == hipSuccess && cur != device_) (void)hipSetDevice(device_);
    if (d_tables_) (void)hipFree(d_tables_);
    if (owns_work_ && d_work_) (void)hipFree(d_work_);
    if (owns_work_ && d_work1_) (void)hipFree(d_work1_);
    if (d_stage_in_) (void)hipFree(d_stage_in_);
    if (d_stage_out_) (void)hipFree(d_stage_out_);
    if (owns_work_ && d_inplace_) (void)hipFree(d_inplace_);
    for (GraphEntry& g : graphs_) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(g.exec));
    delete adjoint_;
    if (capture_stream_) (void)hipStreamDestroy(capture_stream_);
    for (hipStream_t ls : lane_streams_) (void)hipStreamDestroy(ls);
    for (hipEvent_t le : lane_events_) (void)hipEventDestroy(le);
    if (cur >= 0 && device_ >= 0 && cur != device_) (void)hipSetDevice(cur);
}

size_t Executor::work_bytes() const {
    size_t b = align256(work_bytes_) + align256(work1_bytes_);
    if (in_place_unsafe_) b += align256(in_bytes_);
    return std::max(b, twin_work_bytes_);
}

void Executor::inherit_workspace(const Executor& old) {
    auto_alloc_ = old.auto_alloc_;
    if (old.user_work_ != nullptr && work_bytes() <= old.work_bytes()) (void)set_work_area(old.user_work_);
}

void Executor::drop_graphs() {
    for (GraphEntry& g : graphs_) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(g.exec));
    graphs_.clear();
}

void Executor::set_auto_allocation(bool on) {
    auto_alloc_ = on;
    delete adjoint_;  // hipFree in its destructor waits for work still using its buffers
    adjoint_ = nullptr;
}

aclfftResult Executor::set_work_area(void* p) {
    if (p == nullptr || !is_device_pointer(p)) return ACLFFT_INVALID_VALUE;
    delete adjoint_;
    adjoint_ = nullptr;
    if (owns_work_) {
        // plan-allocated buffers (if any) are released; hipFree waits for work using them
        if (d_work_) (void)hipFree(d_work_);
        if (d_work1_) (void)hipFree(d_work1_);
        if (d_inplace_) (void)hipFree(d_inplace_);
    }
    char* base = static_cast<char*>(p);
    user_work_ = base;
    d_work_ = work_bytes_ > 0 ? base : nullptr;
    d_work1_ = work1_bytes_ > 0 ? base + align256(work_bytes_) : nullptr;
    d_inplace_ = in_place_unsafe_ ? base + align256(work_bytes_) + align256(work1_bytes_) : nullptr;
    owns_work_ = false;
    drop_graphs();  // captured launches hold the old addresses
    return ACLFFT_SUCCESS;
}

std::string Executor::describe() const {
    std::ostringstream os;
    const auto& reg = registry();
    for (size_t i = 0; i < steps_.size(); ++i) {
        const Step& s = steps_[i];
        if (i) os << " | ";
        if (s.lane > 0) os << "[lane " << s.lane << "] ";
        if (s.type == Step::COPY) {
            os << "copy";
            continue;
        }
        if (s.type == Step::R2C_UNPACK || s.type == Step::C2R_PACK) {
            os << (s.type == Step::R2C_UNPACK ? "r2c_unpack" : "c2r_pack") << " M=" << s.m;
            continue;
        }
        if (s.type == Step::FUSED2D) {
            const ofx_2d_kernel_info& k = reg_2d()[s.kernel];
            os << "fused2d " << k.nx << "x" << k.ny << " rows=";
            for (int p = 0; p < k.row_npass; ++p) os << (p ? "x" : "") << k.row_radices[p];
            os << " cols=";
            for (int p = 0; p < k.col_npass; ++p) os << (p ? "x" : "") << k.col_radices[p];
            if (s.args.flags & OFX_FLAG_NT_LOAD) os << " ntload";
            if (s.args.flags & OFX_FLAG_NT_STORE) os << " ntstore";
            continue;
        }
        if (s.type == Step::COLUMN) {
            const ofx_kernel_info& k = col_registry()[s.kernel];
            os << (k.kind == OFX_KIND_COL_Z ? "column-z n=" : "column n=") << k.n << " T=" << k.threads_per_transform << " cols=" << k.transforms_per_block
               << " radices=";
            for (int p = 0; p < k.npass; ++p) os << (p ? "x" : "") << k.radices[p];
            if (s.args.ft_len) os << ((s.args.flags & OFX_FLAG_FT_FP64) ? " fourstep-twiddle-fp64" : " fourstep-twiddle");
            continue;
        }
        if (s.type == Step::PERSISTENT) {
            os << "persistent tiles=" << s.b.ntiles << " set=" << s.b.class_set << " sched=" << s.b.schedule
               << (s.b.stream ? " stream" : "") << " prologue=" << s.b.prologue_rounds;
            if (s.blocks_per_cu > 0) os << " bpc=" << s.blocks_per_cu;
            os << " classes=";
            bool first = true;
            for (int c = 0; c < OFX_BCLASS_COUNT; ++c)
                if (s.class_tables[c] >= 0) {
                    ofx_batch_class_info ci;
                    ofx_batch_class_info_get_set(s.b.class_set, c, &ci);
                    os << (first ? "" : ",") << ci.n;
                    first = false;
                }
            continue;
        }
        if (s.type == Step::MFMA) {
            ofx_mfma_info mi{};
            ofx_mfma_kernel_info_get(s.kernel, &mi);
            os << "mfma n=" << s.n << " " << mi.r1 << "x" << mi.r2 << " shape=" << mi.shape << (mi.tiled ? " tiled" : "")
               << " mode=" << (s.mfma_mode == OFX_MFMA_BF16 ? "bf16" : s.mfma_mode == OFX_MFMA_FP16X2 ? "fp16x2" : "bf16x3");
            if (s.tiles_table >= 0) os << " tiles=" << s.args.batch;
            if (s.args.flags & OFX_FLAG_MFMA_2D) os << " 2d=" << mi.r2 << "x" << mi.r1;
            continue;
        }
        if (s.type == Step::GENERIC) {
            os << "generic n=" << s.g.n << " tpb=" << s.g.tpb << " radices=";
            for (int p = 0; p < s.g.npass; ++p) os << (p ? "x" : "") << s.g.radices[p];
            if (s.g.ft_len) os << " fourstep-twiddle";
            continue;
        }
        const ofx_kernel_info& k = reg[s.kernel];
        os << "kind" << k.kind << " n=" << k.n << " T=" << k.threads_per_transform << " TPB=" << k.transforms_per_block
           << (k.iov == 2 ? " paired" : k.iov == 3 ? " hpair" : (k.staged ? " staged" : " direct"))
           << (k.stream == 1 ? " stream" : k.stream == 2 ? " ntstore" : k.stream == 3 ? " ntload" : "")
           << " radices=";
        for (int p = 0; p < k.npass; ++p) os << (p ? "x" : "") << k.radices[p];
        os << " id=" << s.kernel;
    }
    return os.str();
}

// -------------------------------------------------------------------------------------------
// device state
// -------------------------------------------------------------------------------------------
aclfftResult Executor::prepare_device(hipStream_t s) {
    int dev = 0;
    if (hip_log(hipGetDevice(&dev), "hipGetDevice") != hipSuccess) return ACLFFT_SETUP_FAILED;
    if (device_ >= 0) return dev == device_ ? ACLFFT_SUCCESS : ACLFFT_INVALID_DEVICE;
    if (!plan_checked_ && has_plan_alternatives()) {
        // plan-level choice: measured now (ACLFFT_TUNE=1 or aclfftExtTune), else from the cache
        plan_checked_ = true;
        if (force_tune_ || env_int("ACLFFT_TUNE", 0) != 0) {
            aclfftResult r = tune_plan(s);
            if (r != ACLFFT_SUCCESS) return r;
        } else {
            PlanDesc k;
            Executor* w = nullptr;
            if (load_plan_choice(desc_, &k) && build(k, &w) == ACLFFT_SUCCESS) {
                if (owns_work_ || w->work_bytes() <= work_bytes())
                    adopt(w);
                else
                    delete w;
            }
        }
    }
    plan_checked_ = true;
    if (owns_work_ && !auto_alloc_ && (work_bytes_ > 0 || work1_bytes_ > 0))
        return ACLFFT_NO_WORKSPACE;  // caller-managed workspace not set yet
    if (table_bytes_ > 0 && hipMalloc(&d_tables_, table_bytes_) != hipSuccess) {
        (void)hipGetLastError();
        return ACLFFT_ALLOC_FAILED;
    }
    if (owns_work_ && work_bytes_ > 0 && hipMalloc(&d_work_, work_bytes_) != hipSuccess) {
        (void)hipGetLastError();
        return ACLFFT_ALLOC_FAILED;
    }
    if (owns_work_ && work1_bytes_ > 0 && hipMalloc(&d_work1_, work1_bytes_) != hipSuccess) {
        (void)hipGetLastError();
        return ACLFFT_ALLOC_FAILED;
    }
    {
        const aclfftResult rt = gen_tables(d_tables_, s);
        if (rt != ACLFFT_SUCCESS) return rt;
    }
    if (hip_log(hipStreamSynchronize(s), "hipStreamSynchronize (plan setup)") != hipSuccess) return ACLFFT_SETUP_FAILED;
    int lanes = 0;
    for (const Step& st : steps_) lanes = std::max(lanes, st.lane);
    for (int k = 0; k < lanes; ++k) {
        hipStream_t ls = nullptr;
        if (hip_log(hipStreamCreateWithFlags(&ls, hipStreamNonBlocking), "hipStreamCreate (lane)") != hipSuccess)
            return ACLFFT_SETUP_FAILED;
        lane_streams_.push_back(ls);
    }
    for (int k = 0; lanes > 0 && k <= lanes; ++k) {
        hipEvent_t ev = nullptr;
        if (hip_log(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate (lane)") != hipSuccess)
            return ACLFFT_SETUP_FAILED;
        lane_events_.push_back(ev);
    }
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            num_cus_ = cus;
        else
            (void)hipGetLastError();
    }
    device_ = dev;
    if (!tuned_ && (force_tune_ || env_int("ACLFFT_TUNE", 0) != 0)) {
        aclfftResult r = autotune(s);
        if (r != ACLFFT_SUCCESS) return r;
    } else if (!tuned_) {
        apply_tune_cache(desc_, steps_, Registries{registry().data(), (int)registry().size(), col_registry().data(),
                                                   (int)col_registry().size()});
    }
    tuned_ = true;
    return ACLFFT_SUCCESS;
}

aclfftResult Executor::gen_tables(void* base, hipStream_t s) {
    for (const TableSpec& t : tables_) {
        void* p = static_cast<char*>(base) + t.offset;
        int rc = 0;
        if (t.kind == TableSpec::PASS)
            rc = ofx_gen_pass_twiddles(p, t.npass, t.radices, s);
        else if (t.kind == TableSpec::ROOTS)
            rc = ofx_gen_roots(p, t.len, t.count, s);
        else if (t.kind == TableSpec::MFMA_CONST)
            rc = ofx_gen_mfma_kernel_consts(t.radices[0], t.npass, p, s);
        else if (t.kind == TableSpec::PASS_Z)
            rc = ofx_gen_pass_twiddles_z(p, t.npass, t.radices, s);
        else if (t.kind == TableSpec::ROOTS_Z)
            rc = ofx_gen_roots_z(p, t.len, t.count, s);
        else if (t.kind == TableSpec::BLOB)
            rc = (int)hipMemcpyAsync(p, t.blob.data(), t.bytes, hipMemcpyHostToDevice, s);
        else
            rc = ofx_gen_ft_table(p, t.len, s);
        if (rc != 0) {
            (void)hip_log((hipError_t)rc, "twiddle-table generation");
            return ACLFFT_SETUP_FAILED;
        }
    }
    return ACLFFT_SUCCESS;
}

aclfftResult Executor::run(const void* din, void* dout, int direction, hipStream_t s) {
    return run_on(din, dout, direction, s, d_tables_, d_work_, d_work1_);
}

aclfftResult Executor::exec_noalloc(const void* in, void* out, int direction, hipStream_t s) {
    if (!is_device_pointer(in) || !is_device_pointer(out)) return ACLFFT_INVALID_VALUE;
    for (const TableSpec& t : tables_)
        if (t.kind == TableSpec::BLOB) return ACLFFT_NOT_SUPPORTED;  // host-side tables (tile lists)
    for (const Step& st : steps_)
        if (st.lane > 0 || st.type == Step::PERSISTENT) return ACLFFT_NOT_SUPPORTED;
    const bool inplace = in == out && in_place_unsafe_;
    // layout: [tables | work0 | work1 | the in-place copy], 256-byte aligned parts
    const size_t need = align256(table_bytes_) + align256(work_bytes_) + align256(work1_bytes_) +
                        (inplace ? align256(in_bytes_) : 0);
    void* base = nullptr;
    size_t cap = 0;
    if (hip_log((hipError_t)ofx_scratch(&base, &cap), "device scratch lookup") != hipSuccess) return ACLFFT_SETUP_FAILED;
    if (need > cap) {
        std::fprintf(stderr, "[ops-fft] aclfftExtExecC2CNoAlloc: the plan needs %zu bytes of device scratch, "
                             "the library holds %zu per device\n", need, cap);
        return ACLFFT_INVALID_SIZE;
    }
    char* b = static_cast<char*>(base);
    char* tables = b;
    char* w0 = b + align256(table_bytes_);
    char* w1 = w0 + align256(work_bytes_);
    char* cp = w1 + align256(work1_bytes_);
    aclfftResult r = gen_tables(tables, s);
    if (r != ACLFFT_SUCCESS) return r;
    if (inplace) {
        if (hipMemcpyAsync(cp, in, in_bytes_, hipMemcpyDeviceToDevice, s) != hipSuccess) return ACLFFT_EXEC_FAILED;
        in = cp;
    }
    int cus = 0;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cus > 0)
        num_cus_ = cus;
    else
        (void)hipGetLastError();
    return run_on(in, out, direction, s, tables, work_bytes_ > 0 ? w0 : nullptr, work1_bytes_ > 0 ? w1 : nullptr);
}

aclfftResult Executor::run_on(const void* din, void* dout, int direction, hipStream_t s, void* tables, void* work0,
                              void* work1) {
    void* bufs[4] = {const_cast<void*>(din), dout, work0, work1};
    const float dsign = (direction == ACLFFT_BACKWARD) ? -1.f : 1.f;
    auto table = [&](int idx) -> const void* {
        return idx >= 0 ? static_cast<const void*>(static_cast<char*>(tables) + tables_[idx].offset) : nullptr;
    };
    // fork / join of the side-stream lanes: the side streams wait for everything issued so far on
    // `s`, and `s` waits for every side stream before the next lane-0 step and at the end (under
    // stream capture these become the parallel branches of the graph)
    const hipStream_t s0 = s;
    bool forked = false;
    auto fork = [&]() -> bool {
        if (hip_log(hipEventRecord(lane_events_[0], s0), "hipEventRecord (fork)") != hipSuccess) return false;
        for (hipStream_t ls : lane_streams_)
            if (hip_log(hipStreamWaitEvent(ls, lane_events_[0], 0), "hipStreamWaitEvent (fork)") != hipSuccess)
                return false;
        return true;
    };
    auto join = [&]() -> bool {
        for (size_t k = 0; k < lane_streams_.size(); ++k) {
            if (hip_log(hipEventRecord(lane_events_[k + 1], lane_streams_[k]), "hipEventRecord (join)") != hipSuccess ||
                hip_log(hipStreamWaitEvent(s0, lane_events_[k + 1], 0), "hipStreamWaitEvent (join)") != hipSuccess)
                return false;
        }
        return true;
    };
    for (const Step& st : steps_) {
        const char* src = static_cast<const char*>(bufs[st.src]) + st.src_off;
        char* dst = static_cast<char*>(bufs[st.dst]) + st.dst_off;
        hipError_t err = hipSuccess;
        if (st.lane > 0 && (size_t)st.lane <= lane_streams_.size()) {
            if (!forked && !fork()) return ACLFFT_EXEC_FAILED;
            forked = true;
            s = lane_streams_[st.lane - 1];
        } else {
            if (forked && !join()) return ACLFFT_EXEC_FAILED;
            forked = false;
            s = s0;
        }
        switch (st.type) {
            case Step::COPY:
                if (src != dst) err = hipMemcpyAsync(dst, src, st.copy_bytes, hipMemcpyDeviceToDevice, s);
                break;
            case Step::R2C_UNPACK:
                err = st.kind == OFX_KIND_D2Z
                          ? (hipError_t)ofx_r2c_unpack_z(src, dst, table(st.aux_table), st.m, st.args.batch,
                                                         st.args.scale64 != 0.0 ? st.args.scale64 : st.args.scale,
                                                         st.args.flags, s)
                          : (hipError_t)ofx_r2c_unpack(src, dst, table(st.aux_table), st.m, st.args.batch,
                                                       st.args.scale, st.args.flags, s);
                break;
            case Step::C2R_PACK:
                err = st.kind == OFX_KIND_Z2D
                          ? (hipError_t)ofx_c2r_pack_z(src, dst, table(st.aux_table), st.m, st.args.batch, st.args.flags, s)
                          : (hipError_t)ofx_c2r_pack(src, dst, table(st.aux_table), st.m, st.args.batch, st.args.flags, s);
                break;
            case Step::COLUMN:
            case Step::FUSED2D: {
                ofx_args a = st.args;
                a.in = src;
                a.out = dst;
                a.tw = table(st.tw_table);
                a.tw_aux = table(st.aux_table);
                if (st.conj_in) a.in_sign *= dsign;
                if (st.conj_out) a.out_sign *= dsign;
                err = (hipError_t)(st.type == Step::COLUMN ? ofx_launch_col(st.kernel, &a, s)
                                                           : ofx_launch_2d(st.kernel, &a, s));
                break;
            }
            case Step::PERSISTENT: {
                ofx_batch_args b = st.b;
                b.in = src;
                b.out = dst;
                b.tw = tables;
                for (int c = 0; c < OFX_BCLASS_COUNT; ++c)
                    b.tw_off[c] = st.class_tables[c] >= 0 ? (int)(tables_[st.class_tables[c]].offset / 8) : 0;
                b.tiles = static_cast<const ofx_batch_tile*>(table(st.tiles_table));
                b.counter = static_cast<unsigned int*>(const_cast<void*>(table(st.counter_table)));
                b.sign = dsign;
                // by default every workgroup resident at once (4 per CU: 34 KiB of LDS each)
                const int per_cu = st.blocks_per_cu > 0
                                       ? st.blocks_per_cu
                                       : std::max(1, ofx_batch_blocks_per_cu_stream(b.class_set, b.schedule, b.stream));
                err = (hipError_t)ofx_launch_batch_persistent(&b, num_cus_ * per_cu, s);
                break;
            }
            case Step::MFMA: {
                ofx_args a = st.args;
                a.in = src;
                a.out = dst;
                a.tw = table(st.tw_table);
                a.tw_aux = table(st.aux_table);
                if (st.conj_in) a.in_sign *= dsign;
                if (st.conj_out) a.out_sign *= dsign;
                if (st.tiles_table >= 0) a.tiles = table(st.tiles_table);
                err = (hipError_t)ofx_launch_mfma_kernel(st.kernel, st.mfma_mode, &a, env_int("ACLFFT_MFMA_BLOCKS", 0), s);
                break;
            }
            case Step::GENERIC: {
                ofx_args a = st.args;
                a.in = src;
                a.out = dst;
                a.tw = table(st.tw_table);
                a.tw_aux = table(st.aux_table);
                if (st.conj_in) a.in_sign *= dsign;
                if (st.conj_out) a.out_sign *= dsign;
                err = (hipError_t)ofx_launch_generic(&a, &st.g, s);
                break;
            }
            case Step::KERNEL: {
                int kid = st.kernel;
                int twt = st.tw_table;
                if (registry()[kid].staged &&
                    ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15)) {
                    // staged tiles move 16-byte vectors: fall back to a direct-I/O candidate
                    for (size_t c = 0; c < st.candidates.size(); ++c)
                        if (!registry()[st.candidates[c]].staged) {
                            kid = st.candidates[c];
                            twt = st.candidate_tables[c];
                            break;
                        }
                    if (registry()[kid].staged) return ACLFFT_UNALIGNED_DATA;
                }
                ofx_args a = st.args;
                a.in = src;
                a.out = dst;
                a.tw = table(twt);
                a.tw_aux = table(st.aux_table);
                if (st.tiles_table >= 0) a.tiles = table(st.tiles_table);
                if (st.conj_in) a.in_sign *= dsign;
                if (st.conj_out) a.out_sign *= dsign;
                err = (hipError_t)ofx_launch(kid, &a, s);
                break;
            }
        }
        if (hip_log(err, "kernel launch") != hipSuccess) {
            if (forked) (void)join();
            return ACLFFT_EXEC_FAILED;
        }
    }
    if (forked && !join()) return ACLFFT_EXEC_FAILED;
    return ACLFFT_SUCCESS;
}

aclfftResult Executor::exec(const void* in, void* out, int direction, hipStream_t s) {
    aclfftResult r = prepare_device(s);
    if (r != ACLFFT_SUCCESS) return r;
    const bool din = is_device_pointer(in);
    const bool dout = is_device_pointer(out);
    if (din && dout) {
        if (in == out && in_place_unsafe_) {
            // the kernel's output layout overlaps other transforms' inputs: read from a copy
            if (d_inplace_ == nullptr && !auto_alloc_) return ACLFFT_NO_WORKSPACE;
            if (d_inplace_ == nullptr && hipMalloc(&d_inplace_, in_bytes_) != hipSuccess) {
                (void)hipGetLastError();
                return ACLFFT_ALLOC_FAILED;
            }
            if (hipMemcpyAsync(d_inplace_, in, in_bytes_, hipMemcpyDeviceToDevice, s) != hipSuccess)
                return ACLFFT_EXEC_FAILED;
            return run(d_inplace_, out, direction, s);
        }
        if (use_graphs_) return run_graph(in, out, direction, s);
        return run(in, out, direction, s);
    }
    // Host pointers: the reference's synchronous staging semantics
    // (docs/zh/API_Reference/FFT_Common_Interface.md:5).
    const void* src = in;
    void* dst = out;
    if (!din) {
        if (d_stage_in_ == nullptr && hipMalloc(&d_stage_in_, in_bytes_) != hipSuccess) {
            (void)hipGetLastError();
            return ACLFFT_ALLOC_FAILED;
        }
        if (hipMemcpyAsync(d_stage_in_, in, in_bytes_, hipMemcpyHostToDevice, s) != hipSuccess)
            return ACLFFT_EXEC_FAILED;
        src = d_stage_in_;
    }
    if (!dout) {
        if (d_stage_out_ == nullptr && hipMalloc(&d_stage_out_, out_bytes_) != hipSuccess) {
            (void)hipGetLastError();
            return ACLFFT_ALLOC_FAILED;
        }
        dst = d_stage_out_;
    } else if (dst == in) {
        return ACLFFT_INVALID_VALUE;
    }
    r = run(src, dst, direction, s);
    if (r != ACLFFT_SUCCESS) return r;
    if (!dout && hipMemcpyAsync(out, dst, out_bytes_, hipMemcpyDeviceToHost, s) != hipSuccess)
        return ACLFFT_EXEC_FAILED;
    if (hip_log(hipStreamSynchronize(s), "hipStreamSynchronize") != hipSuccess) return ACLFFT_EXEC_FAILED;
    return ACLFFT_SUCCESS;
}

// -------------------------------------------------------------------------------------------
// hipGraph per plan: the launch sequence for (in, out, direction) is captured once on a private
// stream and replayed with hipGraphLaunch on the caller's stream.  The sequence contains kernel and
// copy nodes only (the persistent kernel's work counter is reset by a kernel, not a memset).
// -------------------------------------------------------------------------------------------
aclfftResult Executor::run_graph(const void* din, void* dout, int direction, hipStream_t s) {
    for (const GraphEntry& g : graphs_) {
        if (g.in == din && g.out == dout && g.direction == direction) {
            if (hip_log(hipGraphLaunch(static_cast<hipGraphExec_t>(g.exec), s), "hipGraphLaunch") != hipSuccess)
                return ACLFFT_EXEC_FAILED;
            return ACLFFT_SUCCESS;
        }
    }
    if (capture_stream_ == nullptr &&
        hip_log(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking), "hipStreamCreate") != hipSuccess)
        return ACLFFT_EXEC_FAILED;
    if (hip_log(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture") !=
        hipSuccess)
        return ACLFFT_EXEC_FAILED;
    aclfftResult r = run(din, dout, direction, capture_stream_);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(capture_stream_, &graph);
    if (r != ACLFFT_SUCCESS || hip_log(e, "hipStreamEndCapture") != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return r != ACLFFT_SUCCESS ? r : ACLFFT_EXEC_FAILED;
    }
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (hip_log(e, "hipGraphInstantiate") != hipSuccess) return ACLFFT_EXEC_FAILED;
    // move the executable graph's resources to the device now, not on its first launch
    if (hip_log(hipGraphUpload(ge, s), "hipGraphUpload") != hipSuccess) {
        (void)hipGraphExecDestroy(ge);
        return ACLFFT_EXEC_FAILED;
    }
    if (graphs_.size() >= 8) {  // bounded cache: drop the oldest
        (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(graphs_.front().exec));
        graphs_.erase(graphs_.begin());
    }
    graphs_.push_back(GraphEntry{din, dout, direction, ge});
    if (hip_log(hipGraphLaunch(ge, s), "hipGraphLaunch") != hipSuccess) return ACLFFT_EXEC_FAILED;
    return ACLFFT_SUCCESS;
}

// -------------------------------------------------------------------------------------------
// plan-level tuning: alternative launch lists of the same plan, built host-only, timed on scratch
// buffers; the fastest (among the accurate ones) replaces this executor's launch list
// -------------------------------------------------------------------------------------------
bool Executor::has_plan_alternatives() const { return desc_.uneven || desc_.auto_precision; }

void Executor::release_device() {
    drop_graphs();
    if (d_tables_) (void)hipFree(d_tables_);
    d_tables_ = nullptr;
    if (owns_work_) {
        if (d_work_) (void)hipFree(d_work_);
        if (d_work1_) (void)hipFree(d_work1_);
        if (d_inplace_) (void)hipFree(d_inplace_);
        d_work_ = d_work1_ = d_inplace_ = nullptr;
    }
    if (d_stage_in_) (void)hipFree(d_stage_in_);
    if (d_stage_out_) (void)hipFree(d_stage_out_);
    d_stage_in_ = d_stage_out_ = nullptr;
    for (hipStream_t ls : lane_streams_) (void)hipStreamDestroy(ls);
    for (hipEvent_t le : lane_events_) (void)hipEventDestroy(le);
    lane_streams_.clear();
    lane_events_.clear();
    delete adjoint_;
    adjoint_ = nullptr;
    device_ = -1;
    tuned_ = false;
}

void Executor::adopt(Executor* w) {
    release_device();
    desc_ = w->desc_;
    steps_.swap(w->steps_);
    tables_.swap(w->tables_);
    table_bytes_ = w->table_bytes_;
    work_bytes_ = w->work_bytes_;
    work1_bytes_ = w->work1_bytes_;
    in_bytes_ = w->in_bytes_;
    out_bytes_ = w->out_bytes_;
    in_place_unsafe_ = w->in_place_unsafe_;
    twin_work_bytes_ = w->twin_work_bytes_;
    delete w;
    if (!owns_work_ && user_work_ != nullptr) {
        // the caller's area keeps serving the plan (candidates needing more were not considered)
        char* area = user_work_;
        owns_work_ = true;
        d_work_ = d_work1_ = d_inplace_ = nullptr;
        (void)set_work_area(area);
    }
}

namespace {

// deterministic pseudo-random fill of the tuning input, values in [-1, 1)
void fill_probe(std::vector<float>& h) {
    unsigned long long x = 0x9E3779B97F4A7C15ULL;
    for (float& v : h) {
        x = x * 6364136223846793005ULL + 1442695040888963407ULL;
        v = (float)((double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    }
}

double rel_l2_host(const std::vector<float>& a, const std::vector<float>& ref) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size() && i < ref.size(); ++i) {
        const double d = (double)a[i] - (double)ref[i];
        num += d * d;
        den += (double)ref[i] * ref[i];
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

}  // namespace

aclfftResult Executor::tune_plan(hipStream_t s) {
    if (!has_plan_alternatives()) return ACLFFT_SUCCESS;
    // candidate descriptions
    std::vector<PlanDesc> cands;
    std::vector<int> precs;
    if (desc_.auto_precision) {
        precs.push_back(0);
        for (int m : {OFX_MFMA_BF16, OFX_MFMA_FP16X2, OFX_MFMA_BF16X3}) precs.push_back(m);
    } else {
        precs.push_back(desc_.precision);
    }
    for (int prec : precs) {
        PlanDesc c = desc_;
        c.precision = prec;
        c.uneven_form = 0;
        c.sched = -1;
        c.blocks_per_cu = 0;
        c.class_set = 0;
        c.stream = -1;
        c.prologue = -1;
        if (!desc_.uneven) {
            cands.push_back(c);
            continue;
        }
        // ACLFFT_UNEVEN_MODE=1|2 restricts the comparison to that form
        const int env_form = env_int("ACLFFT_UNEVEN_MODE", 0);
        if (prec == 0 && env_form != 1) {
            PlanDesc t = c;
            t.uneven_form = 2;
            cands.push_back(t);
        }
        if (env_form == 2 && prec == 0) continue;
        for (int set = 0; set < ofx_batch_class_set_count(); ++set)
            for (int sched = 0; sched <= 1; ++sched) {
                const int maxb = std::max(1, ofx_batch_blocks_per_cu_set(set, sched));
                for (int bpc = 1; bpc <= maxb; ++bpc)
                    for (int stream = 0; stream <= 1; ++stream)
                        for (int pro : {3, 0}) {
                            // MFMA precisions: the FP32 part (lengths without MFMA kernels) at
                            // its defaults
                            if (prec != 0 && !(sched == 1 && bpc == maxb && stream == 0 && pro == 3)) continue;
                            PlanDesc t = c;
                            t.uneven_form = 1;
                            t.class_set = set;
                            t.sched = sched;
                            t.blocks_per_cu = bpc;
                            t.stream = prec != 0 ? -1 : stream;
                            t.prologue = prec != 0 ? -1 : pro;
                            cands.push_back(t);
                        }
            }
    }
    // scratch input (pseudo-random) and output, sized for the plan's buffers
    const size_t ib = std::max<size_t>(in_bytes_, 16), ob = std::max<size_t>(out_bytes_, 16);
    void* din = nullptr;
    void* dout = nullptr;
    if (hipMalloc(&din, ib) != hipSuccess || hipMalloc(&dout, ob) != hipSuccess) {
        (void)hipGetLastError();
        if (din) (void)hipFree(din);
        return ACLFFT_ALLOC_FAILED;
    }
    std::vector<float> host(ib / 4);
    fill_probe(host);
    aclfftResult status = ACLFFT_SUCCESS;
    if (hipMemcpy(din, host.data(), ib, hipMemcpyHostToDevice) != hipSuccess) status = ACLFFT_EXEC_FAILED;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    std::vector<float> ref, got;  // the FP32 candidate's output and a candidate's, for the accuracy gate
    const PlanDesc* best = nullptr;
    float best_ms = 1e30f;
    const size_t my_work = work_bytes();
    for (const PlanDesc& c : cands) {
        if (status != ACLFFT_SUCCESS) break;
        Executor* x = nullptr;
        if (build(c, &x) != ACLFFT_SUCCESS) continue;
        if (!owns_work_ && x->work_bytes() > my_work) {  // must fit the caller's work area
            delete x;
            continue;
        }
        x->plan_checked_ = true;
        x->tuned_ = true;  // its step kernels stay at their defaults while plans are compared
        x->use_graphs_ = false;
        bool ok = x->exec(din, dout, ACLFFT_FORWARD, s) == ACLFFT_SUCCESS &&
                  x->exec(din, dout, ACLFFT_FORWARD, s) == ACLFFT_SUCCESS && hipStreamSynchronize(s) == hipSuccess;
        float ms = 0.f;
        if (ok) {
            (void)hipEventRecord(e0, s);
            for (int r = 0; ok && r < 5; ++r) ok = x->exec(din, dout, ACLFFT_FORWARD, s) == ACLFFT_SUCCESS;
            (void)hipEventRecord(e1, s);
            ok = ok && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        }
        if (ok && desc_.auto_precision) {
            // accuracy gate: an MFMA split is kept only within rel-L2 1e-6 of the FP32 result
            std::vector<float>& dst = c.precision == 0 ? ref : got;
            dst.resize(ob / 4);
            ok = hipMemcpy(dst.data(), dout, ob, hipMemcpyDeviceToHost) == hipSuccess;
            if (ok && c.precision != 0) ok = !ref.empty() && rel_l2_host(got, ref) <= 1e-6;
        }
        delete x;
        if (!ok) {
            (void)hipGetLastError();
            continue;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = &c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(din);
    (void)hipFree(dout);
    if (status != ACLFFT_SUCCESS) return status;
    if (best == nullptr) return ACLFFT_SUCCESS;  // keep the current launch list
    Executor* w = nullptr;
    if (build(*best, &w) != ACLFFT_SUCCESS) return ACLFFT_SUCCESS;
    const bool graphs = use_graphs_;
    adopt(w);
    use_graphs_ = graphs;
    store_plan_choice(desc_);
    return ACLFFT_SUCCESS;
}

aclfftResult Executor::tune(hipStream_t s) {
    release_device();
    plan_checked_ = false;
    force_tune_ = true;
    const aclfftResult r = prepare_device(s);
    force_tune_ = false;
    return r;
}

// -------------------------------------------------------------------------------------------
// autotuning: time every candidate kernel of each step on scratch buffers of the plan's size
// -------------------------------------------------------------------------------------------
aclfftResult Executor::autotune(hipStream_t s) {
    bool any = false;
    for (const Step& st : steps_)
        if ((st.type == Step::KERNEL || st.type == Step::COLUMN || st.type == Step::MFMA) && st.candidates.size() > 1)
            any = true;
    if (!any) return ACLFFT_SUCCESS;
    void* a = nullptr;
    void* b = nullptr;
    // candidates run on scratch copies of whichever buffers their step reads and writes (input,
    // output or a four-step workspace with its padded pitch): size both for the largest
    const size_t bytes = std::max(std::max(in_bytes_, out_bytes_), std::max(work_bytes_, work1_bytes_)) + 256;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipGetLastError();
        if (a) (void)hipFree(a);
        return ACLFFT_ALLOC_FAILED;
    }
    (void)hipMemsetAsync(a, 0, bytes, s);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (Step& st : steps_) {
        if ((st.type != Step::KERNEL && st.type != Step::COLUMN && st.type != Step::MFMA) || st.candidates.size() < 2)
            continue;
        auto launch_cand = [&](int id, const ofx_args* a) {
            if (st.type == Step::MFMA) return ofx_launch_mfma_kernel(id, st.mfma_mode, a, 0, s);
            return st.type == Step::COLUMN ? ofx_launch_col(id, a, s) : ofx_launch(id, a, s);
        };
        float best = 1e30f;
        size_t best_i = 0;
        for (size_t ci = 0; ci < st.candidates.size(); ++ci) {
            ofx_args args = st.args;
            args.in = a;
            args.out = b;
            const int tt = st.candidate_tables[ci];
            args.tw = tt >= 0 ? static_cast<char*>(d_tables_) + tables_[tt].offset : nullptr;
            args.tw_aux = st.aux_table >= 0 ? static_cast<char*>(d_tables_) + tables_[st.aux_table].offset : nullptr;
            if (st.tiles_table >= 0) args.tiles = static_cast<char*>(d_tables_) + tables_[st.tiles_table].offset;
            bool ok = launch_cand(st.candidates[ci], &args) == 0;  // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; ok && r < 5; ++r) ok = launch_cand(st.candidates[ci], &args) == 0;
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ok && ms < best) {
                best = ms;
                best_i = ci;
            }
        }
        st.kernel = st.candidates[best_i];
        st.tw_table = st.candidate_tables[best_i];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    store_tune_cache(desc_, steps_, Registries{registry().data(), (int)registry().size(), col_registry().data(),
                                               (int)col_registry().size()});
    return ACLFFT_SUCCESS;
}

}  // namespace ofx
