// executor.h - plan-owned execution state: the launch list chosen at plan time, the device
// twiddle tables, workspace, host-staging buffers and captured graphs.
//
// The reference builds nothing at plan time and, on every exec, regenerates twiddles on the host,
// allocates 5-11 device buffers, copies, launches and synchronises
// (src/fft1_d/arch35/fft/fft1_d_c2c.cpp:188-251).  Here the launch list is decided once
// (MakePlan), device tables are generated on the GPU once (first execution), and an execution on
// device pointers is a pure asynchronous launch sequence on the plan's stream.
#ifndef OFX_EXECUTOR_H
#define OFX_EXECUTOR_H

#include <cstddef>
#include <string>
#include <vector>
#include "cann_ops_fft.h"
#include "ofx_launch.h"

struct ihipEvent_t;  // hipEvent_t, without pulling the HIP runtime headers into the plan layer

namespace ofx {

enum BufId : int { BUF_IN = 0, BUF_OUT = 1, BUF_WORK0 = 2, BUF_WORK1 = 3 };

// Device lookup tables a plan needs; all live in one allocation.
struct TableSpec {
    enum Kind { PASS, ROOTS, FOURSTEP, MFMA_CONST, BLOB, PASS_Z, ROOTS_Z } kind = PASS;  // *_Z: complex128
    int npass = 0;                 // PASS: radix sequence; MFMA_CONST: npass = mode
    int radices[OFX_MAX_PASSES] = {0};
    long long len = 0, count = 0;  // ROOTS: W_len^k, k < count; FOURSTEP: FP64 table of W_len
    size_t offset = 0;             // bytes from the start of the table allocation
    size_t bytes = 0;
    std::vector<char> blob;        // BLOB: host bytes copied to the device at setup
};

// One launch.
struct Step {
    enum Type { KERNEL, GENERIC, COPY, R2C_UNPACK, C2R_PACK, MFMA, PERSISTENT, COLUMN, FUSED2D } type = KERNEL;
    int kernel = -1;               // registry id (KERNEL; FUSED2D: ofx_2d_kernel_* id)
    int kind = OFX_KIND_C2C;
    int n = 0;                     // complex length the kernel computes (for autotuning)
    ofx_args args{};               // in/out/tw/tw_aux and signs filled at execution
    int src = BUF_IN, dst = BUF_OUT;
    size_t src_off = 0, dst_off = 0;  // byte offsets into the buffers
    int tw_table = -1, aux_table = -1;
    bool conj_in = false;          // C2C: apply the direction's conjugation on load / store
    bool conj_out = false;
    size_t copy_bytes = 0;         // COPY
    ofx_generic_desc g{};          // GENERIC
    long long m = 0;               // R2C_UNPACK / C2R_PACK: complex length M = n/2
    int mfma_mode = 0;             // MFMA: enum ofx_mfma_mode
    ofx_batch_args b{};            // PERSISTENT: launch arguments (pointers filled at execution)
    int class_tables[OFX_BCLASS_COUNT] = {0};  // PERSISTENT: pass table of each class (-1 none)
    int tiles_table = -1, counter_table = -1;
    int blocks_per_cu = 0;         // PERSISTENT: workgroups per CU (0: all that fit at once)
    bool final_out = false;        // uneven plans: the last step writing its group's output region
                                   // (the plan's scale is applied there)
    int lane = 0;                  // 0: the caller's stream; k > 0: forked onto the plan's k-th
                                   // side stream (independent steps run concurrently; joined
                                   // back before the next lane-0 step and at the end)
    std::vector<int> candidates;   // registry ids that can replace `kernel` (autotuning)
    std::vector<int> candidate_tables;  // pass-twiddle table of each candidate
};

struct PlanDesc {
    int rank = 1;
    int n0 = 0, n1 = 0, n2 = 0;    // 1D: n0 = length; 2D: n0 = nx (slow), n1 = ny (fast);
                                   // 3D: n0 = nx (slowest), n1 = ny, n2 = nz (fastest)
    long long batch = 1;
    aclfftType type = ACLFFT_C2C;
    bool vertical = false;         // 1D VERTICAL layout: n0 rows x `batch` columns
    int precision = 0;             // 0: FP32 Stockham; else enum ofx_mfma_mode (MFMA split path)
    bool auto_precision = false;   // the tuner chooses `precision` (FP32 or an MFMA split within
                                   // 1e-6 of the FP32 result); FP32 until tuned
    bool hermitian_weight = false; // adjoint real transforms: OFX_FLAG_HERMITIAN_WEIGHT on real steps
    float scale = 1.f;             // multiplies the output (normalisation), fused in the last step
    bool uneven = false;           // aclfftExtPlanManyUneven: groups of (length, batch)
    std::vector<int> ulen, ubatch;
    // knobs of uneven plans, set by the autotuner (zero / negative: the default rule or the
    // ACLFFT_* environment overrides)
    int uneven_form = 0;           // 1 persistent kernel, 2 one tile launch per length
    int sched = -1;                // persistent schedule: 0 dynamic, 1 static
    int blocks_per_cu = 0;         // persistent workgroups per CU (0: all that fit at once)
    int class_set = -1;            // compiled class set of the persistent kernel (negative:
                                   // ACLFFT_PERSISTENT_CLASS_SET, default 0)
    int stream = -1;               // persistent kernel: 1 streaming accesses for its whole-line
                                   // classes, 0 cached; negative: ACLFFT_PERSISTENT_STREAM, else
                                   // streaming when the plan's input plus output is >= 384 MiB
    int prologue = -1;             // persistent kernel: small-tile prologue in tiles per
                                   // workgroup (ofx_batch_args::prologue_rounds); negative:
                                   // ACLFFT_PERSISTENT_PROLOGUE, else 3
};

class Executor {
public:
    // Host-only: choose kernels and build the launch list.  No device call is made.
    static aclfftResult build(const PlanDesc& d, Executor** out);
    ~Executor();

    // Execute on `in` -> `out` (device or host pointers).  direction: ACLFFT_FORWARD/BACKWARD for
    // C2C; ignored for R2C/C2R.
    aclfftResult exec(const void* in, void* out, int direction, hipStream_t stream);

    // Device workspace of one execution: the four-step buffers, then (real plans) the copy an
    // in-place execution on device memory reads from; 256-byte aligned parts.  For real plans it
    // also covers the adjoint execution, which shares the caller's work area.
    size_t work_bytes() const;
    // Caller-managed workspace (aclfftSetWorkArea / aclfftSetAutoAllocation).  With automatic
    // allocation off and no work area, an execution that needs workspace returns NO_WORKSPACE.
    // Both also drop the adjoint executor (it took the previous workspace settings when it was
    // built), so the next adjoint execution rebuilds it with the current ones.
    aclfftResult set_work_area(void* p);
    void set_auto_allocation(bool on);
    // a rebuilt plan (new precision or scale) keeps the caller's workspace settings; the work
    // area carries over when it is large enough for the new launch list
    void inherit_workspace(const Executor& old);
    size_t in_bytes() const { return in_bytes_; }
    size_t out_bytes() const { return out_bytes_; }
    const std::vector<Step>& steps() const { return steps_; }
    const PlanDesc& desc() const { return desc_; }
    std::string describe() const;

    // Measure the alternatives of this plan on scratch buffers on stream s and keep the fastest:
    // for uneven plans the form (persistent kernel or per-length launches), schedule, workgroups
    // per CU and class set; with automatic precision the operand split (MFMA modes only when their
    // result is within rel-L2 1e-6 of the FP32 path's); for every plan each step's kernel
    // candidates.  Synchronous; allocates temporary buffers; the choices go to the tuning cache,
    // from which later plans of the same shape take them on their first execution.
    aclfftResult tune(hipStream_t s);

    // The same execution with no allocation: the device tables and the workspace are placed in
    // the library's device scratch area (ofx_scratch), the tables regenerated on every call.
    // Device pointers only; plans whose tables and workspace exceed the area -> INVALID_SIZE; plans
    // with side-stream lanes or host-side tables (uneven plans) -> NOT_SUPPORTED.
    aclfftResult exec_noalloc(const void* in, void* out, int direction, hipStream_t stream);

private:
    Executor() = default;
    aclfftResult prepare_device(hipStream_t s);
    aclfftResult run(const void* din, void* dout, int direction, hipStream_t s);
    // run with explicit device state: the table allocation and the two workspace buffers
    aclfftResult run_on(const void* din, void* dout, int direction, hipStream_t s, void* tables, void* work0,
                        void* work1);
    aclfftResult gen_tables(void* base, hipStream_t s);
    aclfftResult autotune(hipStream_t s);
    aclfftResult tune_plan(hipStream_t s);
    bool has_plan_alternatives() const;
    // take over w's launch list (this executor's device state is released first); deletes w
    void adopt(Executor* w);
    void release_device();
    int add_table(const TableSpec& t);
    bool plan_checked_ = false;   // plan-level choice looked up (cache) or measured
    bool force_tune_ = false;

    PlanDesc desc_{};
    std::vector<Step> steps_;
    std::vector<TableSpec> tables_;
    size_t table_bytes_ = 0;
    size_t work_bytes_ = 0;       // four-step intermediate
    size_t work1_bytes_ = 0;      // large real transforms: the packed complex data
    size_t in_bytes_ = 0, out_bytes_ = 0;
    size_t twin_work_bytes_ = 0;  // real plans: workspace of the adjoint execution
    bool in_place_unsafe_ = false;  // in == out must go through the workspace

    int device_ = -1;
    void* d_tables_ = nullptr;
    void* d_work_ = nullptr;
    void* d_work1_ = nullptr;
    void* d_stage_in_ = nullptr;
    void* d_stage_out_ = nullptr;
    void* d_inplace_ = nullptr;
    bool tuned_ = false;
    bool auto_alloc_ = true;      // the plan allocates its workspace on first use
    bool owns_work_ = true;       // d_work_ / d_work1_ / d_inplace_ were allocated here
    char* user_work_ = nullptr;   // the caller's work area, when set
    void drop_graphs();
    int num_cus_ = 256;

    // captured launch sequences, keyed by (in, out, direction)
    struct GraphEntry {
        const void* in;
        void* out;
        int direction;
        void* exec;  // hipGraphExec_t
    };
    bool use_graphs_ = false;
    std::vector<GraphEntry> graphs_;
    hipStream_t capture_stream_ = nullptr;
    std::vector<hipStream_t> lane_streams_;  // side streams of steps with lane > 0
    std::vector<ihipEvent_t*> lane_events_;  // hipEvent_t: [0] = fork, [k] = join of lane k
    Executor* adjoint_ = nullptr;
    aclfftResult run_graph(const void* din, void* dout, int direction, hipStream_t s);

public:
    void set_graphs(bool on) { use_graphs_ = on; }
    // the adjoint of a real plan: the opposite real type with the Hermitian weights, built on first use
    Executor* adjoint();

    bool graphs() const { return use_graphs_; }

    friend struct Builder;
};

}  // namespace ofx

#endif
