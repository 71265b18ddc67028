// executor.cpp - kernel choice, device tables, staging and launch sequences of a plan.
#include "executor.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
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

int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

}  // namespace

// -------------------------------------------------------------------------------------------
// plan construction
// -------------------------------------------------------------------------------------------
struct Builder {
    Executor* e;

    // All registry entries computing `kind` at complex length n, in preference order; staged
    // entries only when the transforms are contiguous.
    std::vector<int> kernels_for(int kind, int n, bool contiguous) {
        std::vector<int> ids;
        const auto& reg = registry();
        for (int i = 0; i < (int)reg.size(); ++i) {
            if (reg[i].kind == kind && reg[i].n == n && (contiguous || !reg[i].staged)) ids.push_back(i);
        }
        return ids;
    }

    int pass_table(const ofx_kernel_info& k) {
        if (k.tw_len <= 0) return -1;
        TableSpec t;
        t.kind = TableSpec::PASS;
        t.npass = k.npass;
        for (int p = 0; p < k.npass; ++p) t.radices[p] = k.radices[p];
        t.bytes = (size_t)k.tw_len * 8;
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

    Step kernel_step(int kind, int n, const std::vector<int>& ids) {
        Step s;
        s.type = Step::KERNEL;
        s.kind = kind;
        s.n = n;
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

    // batch of contiguous C2C transforms of length n
    aclfftResult c2c_rows(long long n, long long batch, int src, int dst) {
        if (n == 1) {
            Step s;
            s.type = Step::COPY;
            s.src = src;
            s.dst = dst;
            s.copy_bytes = (size_t)batch * 8;
            s.conj_in = s.conj_out = false;
            e->steps_.push_back(s);
            return ACLFFT_SUCCESS;
        }
        auto ids = kernels_for(OFX_KIND_C2C, (int)n, true);
        if (ids.empty()) return ACLFFT_NOT_IMPLEMENTED;
        Step s = kernel_step(OFX_KIND_C2C, (int)n, ids);
        s.src = src;
        s.dst = dst;
        s.args.batch = batch;
        s.args.idist = n;
        s.args.odist = n;
        s.conj_in = s.conj_out = true;
        e->steps_.push_back(s);
        return ACLFFT_SUCCESS;
    }

    aclfftResult r2c(long long n, long long batch) {
        const long long h = n / 2 + 1;
        if (n == 1) {  // X[0] = x[0]
            return ACLFFT_NOT_IMPLEMENTED;
        }
        const bool even = (n % 2) == 0;
        auto ids = kernels_for(even ? OFX_KIND_R2C : OFX_KIND_R2C_ODD, (int)(even ? n / 2 : n), true);
        if (ids.empty()) return ACLFFT_NOT_IMPLEMENTED;
        Step s = kernel_step(even ? OFX_KIND_R2C : OFX_KIND_R2C_ODD, (int)(even ? n / 2 : n), ids);
        s.src = BUF_IN;
        s.dst = BUF_OUT;
        s.args.batch = batch;
        s.args.idist = n;
        s.args.odist = h;
        if (even) s.aux_table = roots_table(n, n / 2 + 1);
        e->steps_.push_back(s);
        e->in_place_unsafe_ = true;
        return ACLFFT_SUCCESS;
    }

    aclfftResult c2r(long long n, long long batch) {
        const long long h = n / 2 + 1;
        if (n == 1) return ACLFFT_NOT_IMPLEMENTED;
        const bool even = (n % 2) == 0;
        auto ids = kernels_for(even ? OFX_KIND_C2R : OFX_KIND_C2R_ODD, (int)(even ? n / 2 : n), true);
        if (ids.empty()) return ACLFFT_NOT_IMPLEMENTED;
        Step s = kernel_step(even ? OFX_KIND_C2R : OFX_KIND_C2R_ODD, (int)(even ? n / 2 : n), ids);
        s.src = BUF_IN;
        s.dst = BUF_OUT;
        s.args.batch = batch;
        s.args.idist = h;
        s.args.odist = n;
        if (even) s.aux_table = roots_table(n, n / 2 + 1);
        e->steps_.push_back(s);
        e->in_place_unsafe_ = true;
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
    if (d.rank == 1 && !d.vertical) {
        const long long n = d.n0;
        switch (d.type) {
            case ACLFFT_C2C:
                e->in_bytes_ = e->out_bytes_ = (size_t)n * d.batch * 8;
                r = b.c2c_rows(n, d.batch, BUF_IN, BUF_OUT);
                break;
            case ACLFFT_R2C:
                e->in_bytes_ = (size_t)n * d.batch * 4;
                e->out_bytes_ = (size_t)(n / 2 + 1) * d.batch * 8;
                r = b.r2c(n, d.batch);
                break;
            case ACLFFT_C2R:
                e->in_bytes_ = (size_t)(n / 2 + 1) * d.batch * 8;
                e->out_bytes_ = (size_t)n * d.batch * 4;
                r = b.c2r(n, d.batch);
                break;
            default:
                r = ACLFFT_NOT_SUPPORTED;
        }
    }
    if (r != ACLFFT_SUCCESS) {
        delete e;
        return r;
    }
    *out = e;
    return ACLFFT_SUCCESS;
}

Executor::~Executor() {
    int cur = -1;
    if (device_ >= 0 && hipGetDevice(&cur) == hipSuccess && cur != device_) (void)hipSetDevice(device_);
    if (d_tables_) (void)hipFree(d_tables_);
    if (d_work_) (void)hipFree(d_work_);
    if (d_stage_in_) (void)hipFree(d_stage_in_);
    if (d_stage_out_) (void)hipFree(d_stage_out_);
    if (d_inplace_) (void)hipFree(d_inplace_);
    if (cur >= 0 && device_ >= 0 && cur != device_) (void)hipSetDevice(cur);
}

std::string Executor::describe() const {
    std::ostringstream os;
    const auto& reg = registry();
    for (size_t i = 0; i < steps_.size(); ++i) {
        const Step& s = steps_[i];
        if (i) os << " | ";
        if (s.type == Step::COPY) {
            os << "copy";
            continue;
        }
        const ofx_kernel_info& k = reg[s.kernel];
        os << "kind" << k.kind << " n=" << k.n << " T=" << k.threads_per_transform << " TPB=" << k.transforms_per_block
           << (k.staged ? " staged" : " direct") << " radices=";
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
    if (table_bytes_ > 0 && hipMalloc(&d_tables_, table_bytes_) != hipSuccess) {
        (void)hipGetLastError();
        return ACLFFT_ALLOC_FAILED;
    }
    if (work_bytes_ > 0 && hipMalloc(&d_work_, work_bytes_) != hipSuccess) {
        (void)hipGetLastError();
        return ACLFFT_ALLOC_FAILED;
    }
    for (const TableSpec& t : tables_) {
        void* p = static_cast<char*>(d_tables_) + t.offset;
        int rc = 0;
        if (t.kind == TableSpec::PASS)
            rc = ofx_gen_pass_twiddles(p, t.npass, t.radices, s);
        else
            rc = ofx_gen_roots(p, t.len, t.count, s);
        if (rc != 0) {
            (void)hip_log((hipError_t)rc, "twiddle-table generation");
            return ACLFFT_SETUP_FAILED;
        }
    }
    if (hip_log(hipStreamSynchronize(s), "hipStreamSynchronize (plan setup)") != hipSuccess) return ACLFFT_SETUP_FAILED;
    device_ = dev;
    if (!tuned_ && env_int("ACLFFT_TUNE", 0) != 0) {
        aclfftResult r = autotune(s);
        if (r != ACLFFT_SUCCESS) return r;
    } else if (!tuned_) {
        apply_tune_cache(desc_, steps_, registry().data(), (int)registry().size());
    }
    tuned_ = true;
    return ACLFFT_SUCCESS;
}

aclfftResult Executor::run(const void* din, void* dout, int direction, hipStream_t s) {
    void* bufs[4] = {const_cast<void*>(din), dout, d_work_, d_work_};
    const float dsign = (direction == ACLFFT_BACKWARD) ? -1.f : 1.f;
    for (const Step& st : steps_) {
        const char* src = static_cast<const char*>(bufs[st.src]) + st.src_off;
        char* dst = static_cast<char*>(bufs[st.dst]) + st.dst_off;
        if (st.type == Step::COPY) {
            if (src != dst && hipMemcpyAsync(dst, src, st.copy_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
                return ACLFFT_EXEC_FAILED;
            continue;
        }
        int kid = st.kernel;
        int twt = st.tw_table;
        if (registry()[kid].staged && ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15)) {
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
        a.tw = twt >= 0 ? static_cast<char*>(d_tables_) + tables_[twt].offset : nullptr;
        a.tw_aux = st.aux_table >= 0 ? static_cast<char*>(d_tables_) + tables_[st.aux_table].offset : nullptr;
        if (st.conj_in) a.in_sign *= dsign;
        if (st.conj_out) a.out_sign *= dsign;
        if (hip_log((hipError_t)ofx_launch(kid, &a, s), "kernel launch") != hipSuccess) return ACLFFT_EXEC_FAILED;
    }
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
            if (d_inplace_ == nullptr && hipMalloc(&d_inplace_, in_bytes_) != hipSuccess) {
                (void)hipGetLastError();
                return ACLFFT_ALLOC_FAILED;
            }
            if (hipMemcpyAsync(d_inplace_, in, in_bytes_, hipMemcpyDeviceToDevice, s) != hipSuccess)
                return ACLFFT_EXEC_FAILED;
            return run(d_inplace_, out, direction, s);
        }
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
// autotuning: time every candidate kernel of each step on scratch buffers of the plan's size
// -------------------------------------------------------------------------------------------
aclfftResult Executor::autotune(hipStream_t s) {
    bool any = false;
    for (const Step& st : steps_)
        if (st.type == Step::KERNEL && st.candidates.size() > 1) any = true;
    if (!any) return ACLFFT_SUCCESS;
    void* a = nullptr;
    void* b = nullptr;
    const size_t bytes = std::max(in_bytes_, out_bytes_) + 256;
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
        if (st.type != Step::KERNEL || st.candidates.size() < 2) continue;
        float best = 1e30f;
        size_t best_i = 0;
        for (size_t ci = 0; ci < st.candidates.size(); ++ci) {
            ofx_args args = st.args;
            args.in = a;
            args.out = b;
            const int tt = st.candidate_tables[ci];
            args.tw = tt >= 0 ? static_cast<char*>(d_tables_) + tables_[tt].offset : nullptr;
            args.tw_aux = st.aux_table >= 0 ? static_cast<char*>(d_tables_) + tables_[st.aux_table].offset : nullptr;
            bool ok = ofx_launch(st.candidates[ci], &args, s) == 0;  // warm-up
            (void)hipEventRecord(e0, s);
            for (int r = 0; ok && r < 5; ++r) ok = ofx_launch(st.candidates[ci], &args, s) == 0;
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
    store_tune_cache(desc_, steps_, registry().data());
    return ACLFFT_SUCCESS;
}

}  // namespace ofx
