/*
 * ofx_launch.h - thin C-ABI between the host plan layer (ops-fft_amd/lib) and the HIP kernels
 * (ops-fft_amd/csrc) of libcann_ops_fft.so.
 *
 * This layer replaces the reference's per-operator host entries, which take
 * (float* x, float* y, uint32_t n, int32_t norm, uint32_t batches, int isForward, void* stream)
 * and allocate, stage, launch and synchronise inside every call:
 *   aclfftFft1DC2C / aclfftFft1DC2CMix   (src/fft1_d/fft1_d.h:37-58),
 *   aclfftRfft1DFft                       (src/rfft1_d/rfft1_d.h:47-74),
 *   aclfftIrfft1DFft                      (src/irfft1_d/irfft1_d.h:32-39),
 *   aclfftFft2DDd                         (src/fft2_d/fft2_d.h:27-28).
 * Here the plan is built once (kernel choice, device twiddle tables, workspace) and each call is a
 * pure asynchronous launch on a HIP stream with device pointers; nothing allocates or blocks.
 *
 * Kernels are identified by an index into a registry of compiled configurations; the host queries
 * the registry (ofx_kernel_count / ofx_kernel_info) to choose one.  All functions return a
 * hipError_t value (0 = success).
 */
#ifndef OFX_LAUNCH_H
#define OFX_LAUNCH_H

#include <stdint.h>
#include <stddef.h>

struct ihipStream_t;
typedef struct ihipStream_t* hipStream_t;

#if defined(__GNUC__)
#define OFX_API __attribute__((visibility("default")))
#else
#define OFX_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define OFX_MAX_PASSES 12

/* Kernel families. */
enum ofx_kind {
    OFX_KIND_C2C = 0,      /* batch of contiguous length-n complex transforms             */
    OFX_KIND_R2C = 1,      /* even n: n/2-point complex FFT + fused Hermitian unpack       */
    OFX_KIND_C2R = 2,      /* even n: fused Hermitian pre-pack + n/2-point complex FFT     */
    OFX_KIND_COL = 3,      /* strided "column" transforms (2D second pass, four-step)      */
    OFX_KIND_R2C_ODD = 4,  /* odd n: full complex FFT of real input, first n/2+1 bins      */
    OFX_KIND_C2R_ODD = 5,  /* odd n: Hermitian expansion + full complex FFT, real part     */
    OFX_KIND_ROWT = 6,     /* contiguous rows in, transposed out: transform t -> out +
                              (t / cols)*odist + rhi*odist_hi + rlo + k*ostride with
                              t % cols = rhi*cols_lo + rlo (four-step row step)          */
    OFX_KIND_C2C_TILES = 7,/* C2C over a tile list (uneven batches): workgroup b runs
                              tiles[b] = (offset, count <= TPB) contiguous transforms at
                              element offset `offset` of in and out; batch = tile count  */
    /* complex128 / double counterparts of C2C, R2C, C2R, R2C_ODD, C2R_ODD: same layouts and
       arguments, elements of 16 (complex) / 8 (real) bytes, tables from the *_z generators  */
    OFX_KIND_Z2Z = 8,
    OFX_KIND_D2Z = 9,
    OFX_KIND_Z2D = 10,
    OFX_KIND_D2Z_ODD = 11,
    OFX_KIND_Z2D_ODD = 12,
    OFX_KIND_ROWT_Z = 13,  /* complex128 ROWT                                             */
    OFX_KIND_COL_Z = 14    /* complex128 column kernels (ofx_col_kernel_* registry)          */
};

/* Description of one compiled kernel configuration. */
typedef struct ofx_kernel_info {
    int kind;                       /* enum ofx_kind                                        */
    int n;                          /* complex FFT length the kernel computes               */
    int threads_per_transform;      /* T                                                    */
    int transforms_per_block;       /* TPB (for COL: columns per block)                     */
    int block_threads;              /* T * TPB                                              */
    int npass;                      /* number of Stockham passes                            */
    int radices[OFX_MAX_PASSES];    /* radix of each pass, first to last                    */
    int staged;                     /* 1: 16-byte HBM accesses (staged tiles or paired I/O):
                                       needs 16-byte aligned rows                            */
    int lds_bytes;                  /* static LDS per workgroup                             */
    int tw_len;                     /* complex64 entries of the pass-twiddle table          */
    int iov;                        /* 2: paired butterflies in the first/last pass;
                                       3: R2C / C2R with the Hermitian partner butterflies of
                                       the last / first pass in one thread                  */
    int stream;                     /* 1: streaming (non-temporal) HBM input and output, for
                                       working sets well beyond the last-level cache          */
} ofx_kernel_info;

/* Arguments of every 1D launch.  Element distances are in complex64 units for complex data and
 * in float units for real data. */
typedef struct ofx_args {
    const void* in;
    void* out;
    const void* tw;        /* pass-twiddle table (ofx_gen_pass_twiddles)                  */
    const void* tw_aux;    /* R2C/C2R: W_n^k, k in [0, n/2]; four-step: unused            */
    long long batch;       /* number of transforms                                        */
    long long idist;       /* distance between consecutive transforms in `in`             */
    long long odist;       /* distance between consecutive transforms in `out`            */
    long long istride;     /* COL: distance between consecutive elements of a transform   */
    long long ostride;
    float in_sign;         /* -1 conjugates the input on load (backward = conj o fwd o conj) */
    float out_sign;        /* -1 conjugates the output on store                            */
    float scale;           /* multiplies every output value                                */
    int flags;             /* OFX_FLAG_*                                                   */
    long long cols;        /* COL kernels: columns per image (batch = images * cols)       */
    long long ft_len;      /* COL kernels: > 0 multiplies output i of column c by the four-step
                              twiddle W_ft_len^(c*i) (tw_aux = ofx_gen_ft_table of ft_len)   */
    long long cols_lo;     /* ROWT: splits the row index r = t % cols into r / cols_lo (stride
                              odist_hi in `out`) and r % cols_lo (stride 1); COL: splits the
                              output image index m into m / cols_lo (stride odist_hi) and
                              m % cols_lo (stride odist); 0 = no split                        */
    long long odist_hi;
    const void* tiles;     /* C2C_TILES: ofx_batch_tile array of `batch` entries (device)    */
    double scale64;        /* complex128 kernels: output scale in double (0 = use `scale`)   */
    long long ft_div;      /* COL kernels with ft_len: the twiddle uses column c / ft_div (the
                              four-step of VERTICAL columns, where `cols` interleaves the
                              transforms' columns); 0 or 1 = c                                 */
} ofx_args;

/* Runtime-radix ("generic") launches: any length whose radix sequence is listed, any layout.
 * Transform t is addressed as (t_hi, t_lo) = (t / cnt_lo, t % cnt_lo):
 *   input  element i at in  + t_hi*idist + t_lo*idist_lo + i*istride
 *   output element k at out + t_hi*odist + t_lo*odist_lo + k*ostride
 * (units: complex64 or float as for ofx_args; complex128 or double when desc.z = 1).  All data of a
 * workgroup's transforms is held in LDS (two buffers), so tpb * n <= OFX_GENERIC_MAX_ELEMS
 * (OFX_GENERIC_MAX_ELEMS_Z for complex128). */
#define OFX_GENERIC_MAX_ELEMS 9216
#define OFX_GENERIC_MAX_ELEMS_Z 4608

enum ofx_in_mode {
    OFX_IN_COMPLEX = 0,    /* complex64, conjugated when in_sign = -1                         */
    OFX_IN_REAL = 1,       /* float, imaginary part 0 (odd-n R2C)                              */
    OFX_IN_HERMITIAN = 2,  /* n/2+1 bins expanded by conjugate symmetry (odd-n C2R)            */
    OFX_IN_C2R_PACK = 3    /* even C2R pre-pack of 2n-point spectra into n complex values      */
};
enum ofx_out_mode {
    OFX_OUT_COMPLEX = 0,   /* complex64, conjugated when out_sign = -1                         */
    OFX_OUT_HALF = 1,      /* bins 0..n/2 only (odd-n R2C)                                     */
    OFX_OUT_REAL = 2,      /* real part (odd-n C2R)                                            */
    OFX_OUT_R2C_UNPACK = 3 /* even R2C Hermitian unpack of n complex values into n+1 bins      */
};

typedef struct ofx_generic_desc {
    int n;                          /* complex FFT length                                     */
    int npass;
    int radices[OFX_MAX_PASSES];
    int tw_off[OFX_MAX_PASSES];     /* offset of each pass in the pass-twiddle table          */
    int tpb;                        /* transforms per workgroup, 1 .. 64                       */
    int in_mode, out_mode;
    int in_tmajor, out_tmajor;      /* 1: consecutive lanes walk consecutive transforms (strided
                                       layouts), 0: consecutive elements of one transform      */
    int z;                          /* 1: complex128 data, FP64 tables (tpb * n <=
                                       OFX_GENERIC_MAX_ELEMS_Z); 0: complex64                   */
    long long cnt_lo;               /* transforms per t_hi group (>= 1)                       */
    long long idist_lo, odist_lo;
    long long ft_len;               /* four-step: output k of transform t is multiplied by
                                       W_ft_len^((t_lo + ft_base) * k); 0 = none               */
    long long ft_base;
} ofx_generic_desc;

/* Radix sequences the generic kernel can run: every radix must be one of these. */
OFX_API int ofx_generic_radix_supported(int radix);

OFX_API int ofx_launch_generic(const ofx_args* args, const ofx_generic_desc* desc, hipStream_t stream);

/* FP64 four-step twiddle table for W_len^e (two levels, ofx_ft_table_count(len) double2 entries),
 * generated on the GPU with FP64 sincospi. */
OFX_API long long ofx_ft_table_count(long long len);
OFX_API int ofx_gen_ft_table(void* table, long long len, hipStream_t stream);

/* Large even-n real transforms (n = 2M): Hermitian unpack of M-point spectra Z [batch][M] into
 * X [batch][M+1] (W = W_n^k, k <= M), and the C2R pre-pack X -> conj(Z') [batch][M]. */
OFX_API int ofx_r2c_unpack(const void* Z, void* X, const void* W, long long M, long long batch, float scale,
                           int flags, hipStream_t stream);
OFX_API int ofx_c2r_pack(const void* X, void* Z, const void* W, long long M, long long batch, int flags,
                         hipStream_t stream);
/* The same for complex128 data (D2Z / Z2D beyond the single-workgroup lengths; W in FP64). */
OFX_API int ofx_r2c_unpack_z(const void* Z, void* X, const void* W, long long M, long long batch, double scale,
                             int flags, hipStream_t stream);
OFX_API int ofx_c2r_pack_z(const void* X, void* Z, const void* W, long long M, long long batch, int flags,
                           hipStream_t stream);

/* Streaming device copy (16-byte non-temporal loads and stores), `bytes` a multiple of 16 and both
 * pointers 16-byte aligned: the achievable-bandwidth yardstick the benchmarks measure beside the
 * transforms.  No reference counterpart (a measurement aid). */
OFX_API int ofx_stream_copy(const void* src, void* dst, long long bytes, hipStream_t stream);

/* MFMA split-precision C2C (dense 32-point DFT stages on v_mfma_f32_32x32x16, FP32 accumulate).
 * Operand precision: */
enum ofx_mfma_mode {
    OFX_MFMA_BF16 = 1,    /* one BF16 part per operand (about 3 significant digits)          */
    OFX_MFMA_FP16X2 = 2,  /* FP16 hi + lo parts, 3 products (about FP32 accuracy)            */
    OFX_MFMA_BF16X3 = 3   /* BF16 hi + mid + lo parts, 6 products (about FP32 accuracy)      */
};
/* bytes of the constant-operand table of `mode` and its generation (FP64 sincospi on the GPU) */
OFX_API long long ofx_mfma_const_bytes(int mode);
OFX_API int ofx_gen_mfma_consts(int mode, void* table, hipStream_t stream);
/* Batch of length-n C2C transforms (n = 1024), contiguous (idist/odist in complex64 units).
 * args->tw = constant table of `mode`, args->tw_aux = W_1024^k (k < 1024, ofx_gen_roots).
 * blocks: workgroups of the grid-stride launch (0 = default: 2048, or 512 for OFX_MFMA_BF16X3); a
 * negative value -b launches b
 * workgroups of the software-pipelined variant (next transform's loads issued before the MFMA work
 * of the current one). */
OFX_API int ofx_launch_mfma_c2c(int mode, int n, const ofx_args* args, int blocks, hipStream_t stream);

/* Registry of the MFMA C2C kernels: n = r1 * r2 as two dense DFT stages (r1, r2 in {16, 32, 64}),
 * or, with r1 = 1, n = r2 in {16, 32, 64} as one dense stage for 16 transforms per wave (no tile
 * lists, no 2D form).
 * Entry 0 is the n = 1024 kernel above (shape 32: v_mfma_f32_32x32x16); the others run on
 * v_mfma_f32_16x16x32 (shape 16).  Same arguments as ofx_launch_mfma_c2c: contiguous transforms,
 * args->tw = the entry's constant table for `mode` (ofx_gen_mfma_kernel_consts, bytes from
 * ofx_mfma_kernel_const_bytes), args->tw_aux = W_n^k for k < n (ofx_gen_roots); blocks = grid
 * (0 = default); args->flags OFX_FLAG_MFMA_2D (entries >= 1): 2D images of r2 x r1 (the
 * reference's 2D matrix form P X Q, src/fft2_d/arch32/dd/fft2_d_dd.cpp:31-66).  No other
 * reference counterpart beyond the matrix-DFT stage shapes of its Cube kernels. */
typedef struct ofx_mfma_info {
    int n;
    int r1;      /* stage-2 DFT length (n = n1 + r1 * n2); 1: the single-stage dense kernels      */
    int r2;      /* stage-1 DFT length                                                            */
    int shape;   /* 32 or 16: the MFMA tile                                                       */
    int tiled;   /* 1: the input is read one stage-1 row tile at a time (fewer registers)          */
} ofx_mfma_info;
OFX_API int ofx_mfma_kernel_count(void);
OFX_API int ofx_mfma_kernel_info_get(int id, ofx_mfma_info* info);
OFX_API long long ofx_mfma_kernel_const_bytes(int id, int mode);
OFX_API int ofx_gen_mfma_kernel_consts(int id, int mode, void* table, hipStream_t stream);
/* With args->tiles set (entries other than 0): a tile-list launch for uneven batches, `batch` = the
 * number of ofx_batch_tile entries; workgroup b transforms tiles b, b + grid, ... (grid = blocks,
 * default min(batch, 1024)), each of tiles[i].count contiguous transforms at complex64 offset
 * tiles[i].offset of in and out (offsets may be odd). */
OFX_API int ofx_launch_mfma_kernel(int id, int mode, const ofx_args* args, int blocks, hipStream_t stream);

/* Persistent batched C2C over uneven lengths: compiled length classes. */
enum ofx_batch_class {
    OFX_BCLASS_60 = 0, OFX_BCLASS_128, OFX_BCLASS_243, OFX_BCLASS_256, OFX_BCLASS_512, OFX_BCLASS_1000,
    OFX_BCLASS_1024, OFX_BCLASS_2048, OFX_BCLASS_4096, OFX_BCLASS_6561, OFX_BCLASS_COUNT
};
typedef struct ofx_batch_class_info {
    int n;                          /* transform length                                       */
    int transforms_per_tile;        /* transforms one workgroup processes together            */
    int npass;
    int radices[OFX_MAX_PASSES];
    int tw_len;                     /* entries of the class's pass-twiddle table              */
} ofx_batch_class_info;
typedef struct ofx_batch_tile {
    int cls;                        /* enum ofx_batch_class                                   */
    int count;                      /* transforms in this tile (1 .. transforms_per_tile)     */
    long long offset;               /* complex64 offset of the tile's first transform         */
} ofx_batch_tile;
typedef struct ofx_batch_args {
    const void* in;
    void* out;
    const void* tw;                 /* all classes' pass-twiddle tables                       */
    int tw_off[OFX_BCLASS_COUNT];   /* complex64 offset of each class's table in `tw`         */
    const ofx_batch_tile* tiles;    /* device array, processed in order of the work counter   */
    int ntiles;
    float sign;                     /* +1 forward, -1 backward (conjugation on load/store)    */
    unsigned int* counter;          /* device work counter (reset by the launch; dynamic only)  */
    int schedule;                   /* 0 dynamic: after its first tile (blockIdx), a workgroup
                                       takes the next index from the counter, fetched one tile
                                       ahead; 1 static: the k-th tile of workgroup b is walk
                                       position k*blocks + (k even ? b : blocks-1-b) (the list
                                       is sorted by cost, so the rounds alternate direction),
                                       its descriptor loaded one tile ahead; no counter.  Walk
                                       positions map to list entries as prologue_rounds says   */
    float scale;                    /* multiplies every output value (0 is taken as 1)         */
    int class_set;                  /* compiled class set (ofx_batch_class_set_count; 0 default)  */
    int stream;                     /* 1: streaming (non-temporal) HBM accesses for the classes
                                       whose transforms are whole 128-byte lines (n % 16 == 0),
                                       for working sets well beyond the last-level cache      */
    int prologue_rounds;            /* the walk's first min(prologue_rounds * blocks, ntiles / 2)
                                       positions take the smallest tiles, in increasing cost (the
                                       list's end, backwards); the rest the list from its start  */
} ofx_batch_args;
OFX_API int ofx_batch_class_count(void);
OFX_API int ofx_batch_class_info_get(int cls, ofx_batch_class_info* info);   /* class set 0 */
/* The persistent kernel is compiled for several class sets (different radix orders / threads per
 * transform of some length classes); a launch names its set in ofx_batch_args.class_set. */
OFX_API int ofx_batch_class_set_count(void);
OFX_API int ofx_batch_class_info_get_set(int set, int cls, ofx_batch_class_info* info);
OFX_API int ofx_batch_blocks_per_cu_set(int set, int schedule);
/* The same for the streaming (stream = 1) or cached (0) variant of the kernel. */
OFX_API int ofx_batch_blocks_per_cu_stream(int set, int schedule, int stream);
/* Launches `blocks` persistent workgroups (dynamic schedule: after a one-thread kernel that
 * resets the counter). */
OFX_API int ofx_launch_batch_persistent(const ofx_batch_args* args, int blocks, hipStream_t stream);
/* Workgroups of the persistent kernel that fit on one CU at once (registers, LDS), for the current
 * device; the grid for all of them resident is this times the CU count. */
OFX_API int ofx_batch_blocks_per_cu(int schedule);

/* R2C: multiply bins 1..n/2-1 by 2 (adjoint of C2R).  C2R: multiply input bins 1..n/2-1 by 1/2
 * (adjoint of R2C).  Used by the hand-written adjoint kernels. */
#define OFX_FLAG_HERMITIAN_WEIGHT 1
/* COL kernels with a power-of-two ft_len (complex64): form the four-step twiddle in FP64 (the
 * factored form's per-lane and per-column factors from FP64 sincospi, their product in FP64, one
 * rounding to FP32) instead of from FP32 sincospif factors (about 2 ulp). */
#define OFX_FLAG_FT_FP64 2
/* Streaming (non-temporal) HBM loads / stores, for the launches that take them as flags (the
 * one-pass 2D kernels) */
#define OFX_FLAG_NT_LOAD 4
#define OFX_FLAG_NT_STORE 8
/* One-pass 2D kernels: run the non-pipelined form (one image per workgroup) even where the
 * pipelined one applies */
#define OFX_FLAG_NO_PIPELINE 16
/* MFMA DFT kernels (ofx_launch_mfma_kernel, ids >= 1): 2D images instead of 1D transforms.  Each
 * transform is an image of r2 rows and r1 columns (row-major); with args->tw_aux all ones (W_1^k,
 * ofx_gen_roots with len 1) stage 1 transforms the columns, stage 2 the rows, and the output is
 * stored in the image layout (element (k_row, k_col) at r1 * k_row + k_col). */
#define OFX_FLAG_MFMA_2D 32

/* Device scratch area of the allocation-free execution entry point (aclfftExtExecC2CNoAlloc):
 * OFX_SCRATCH_BYTES on every GPU, a file-scope device array held from the library's load until the
 * process ends.  *ptr = its address on the current device (looked up once per device), *bytes =
 * its size (optional). */
#define OFX_SCRATCH_BYTES ((size_t)64 << 20)
OFX_API int ofx_scratch(void** ptr, size_t* bytes);

OFX_API int ofx_kernel_count(void);
OFX_API int ofx_kernel_info_get(int id, ofx_kernel_info* info);

/* Launch registry kernel `id` on `stream`.  Returns 0 or a hipError_t. */
OFX_API int ofx_launch(int id, const ofx_args* args, hipStream_t stream);

/* Column kernels (OFX_KIND_COL): `batch` = images * cols transforms of length n; column c of image
 * m has element i at in[m*idist + c + i*istride] (out likewise with odist/ostride).
 * transforms_per_block of the info = adjacent columns per workgroup. */
OFX_API int ofx_col_kernel_count(void);
OFX_API int ofx_col_kernel_info_get(int id, ofx_kernel_info* info);
OFX_API int ofx_launch_col(int id, const ofx_args* args, hipStream_t stream);

/* One-pass 2D C2C kernels (complex64) for images that fit in LDS, nx, ny in {32, 64, 128} - the
 * reference's 2D domain (src/fft2_d/arch32/fft2_d_exec_api.cpp:18-60, dd/fft2_d_dd.cpp:17-174):
 * `batch` images of nx*ny complex64 (row-major, ny fastest), image m at in + m*idist and
 * out + m*odist; tw = the row passes' twiddle table (row radices),
 * tw_aux = the column passes' (column radices), both from ofx_gen_pass_twiddles; in_sign
 * conjugates on load, out_sign and scale apply on store.  In place is allowed (each workgroup reads
 * its whole image before writing it). */
typedef struct ofx_2d_kernel_info {
    int nx, ny;                          /* image rows (slow axis), columns (fast axis)        */
    int block_threads;
    int lds_bytes;
    int row_npass, row_radices[OFX_MAX_PASSES], row_tw_len;   /* length-ny row transforms    */
    int col_npass, col_radices[OFX_MAX_PASSES], col_tw_len;   /* length-nx column transforms */
} ofx_2d_kernel_info;
OFX_API int ofx_2d_kernel_count(void);
OFX_API int ofx_2d_kernel_info_get(int id, ofx_2d_kernel_info* info);
OFX_API int ofx_launch_2d(int id, const ofx_args* args, hipStream_t stream);

/* Fill `tw` (device, tw_len complex64) with the pass-twiddle table of a radix sequence:
 * FP64 sincospi on the GPU, rounded once to FP32. */
OFX_API int ofx_gen_pass_twiddles(void* tw, int npass, const int* radices, hipStream_t stream);

/* Fill `tw` (device, count complex64) with W_len^k = e^{-2 pi i k/len}, k = 0..count-1 (FP64
 * sincospi on the GPU, rounded to FP32). */
OFX_API int ofx_gen_roots(void* tw, long long len, long long count, hipStream_t stream);
/* complex128 versions of the two generators above (double2 entries, FP64 values not rounded). */
OFX_API int ofx_gen_pass_twiddles_z(void* tw, int npass, const int* radices, hipStream_t stream);
OFX_API int ofx_gen_roots_z(void* tw, long long len, long long count, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* OFX_LAUNCH_H */
