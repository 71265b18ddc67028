/*
 * ops_fft_ext.h - entry points of libcann_ops_fft.so that have no counterpart in the reference
 * aclfft API.  They are additive: nothing in cann_ops_fft.h changes meaning because of them.
 */
#ifndef OPS_FFT_EXT_H
#define OPS_FFT_EXT_H

#include "cann_ops_fft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Writes a one-line description of the plan's launch sequence (kernel family, lengths, radices)
 * into buf (NUL-terminated, truncated to len).  For tests, tuning and logs. */
ACLFFT_API aclfftResult aclfftExtDescribe(aclfftHandle plan, char* buf, size_t len);

/* Arithmetic of the plan's transforms:
 *   ACLFFT_EXT_PRECISION_FP32     (0) FP32 Stockham kernels (default; the reference's precision)
 *   ACLFFT_EXT_PRECISION_BF16     (1) MFMA dense-DFT stages, one BF16 part per operand
 *   ACLFFT_EXT_PRECISION_FP16X2   (2) MFMA, FP16 hi + lo split operands (about FP32 accuracy)
 *   ACLFFT_EXT_PRECISION_BF16X3   (3) MFMA, BF16 hi + mid + lo split operands
 * The MFMA modes run n = r1 * r2 as two dense DFT stages of 16, 32 or 64 points, and n = 16, 32
 * or 64 as one dense stage (the reference's dense small-n DFT, src/fft1_d/arch32/dft/
 * fft1_d_dft.cpp:23-77): 1D C2C plans of length 16, 32, 64, 256, 512, 1024, 2048 or 4096; 1D R2C /
 * C2R plans of length 32, 64, 128, 512, 1024, 2048, 4096 or 8192
 * (the half-length complex transform on MFMA, the Hermitian unpack / pre-pack in FP32); uneven
 * plans (aclfftExtPlanManyUneven) with at least one such length (its groups run on MFMA, the
 * other lengths stay on the FP32 kernels); 2D C2C plans whose axes are 16, 32 or 64 points (each
 * image as F X F, the column and row DFTs as the two dense stages: the matrix form of the
 * reference's 2D kernel, src/fft2_d/arch32/dd/fft2_d_dd.cpp:31-66; its 128-point axes have no
 * MFMA stage).  Other plans return ACLFFT_NOT_SUPPORTED and keep their current precision.
 * Rebuilds the plan's launch list. */
#define ACLFFT_EXT_PRECISION_FP32 0
#define ACLFFT_EXT_PRECISION_BF16 1
#define ACLFFT_EXT_PRECISION_FP16X2 2
#define ACLFFT_EXT_PRECISION_BF16X3 3
/*   ACLFFT_EXT_PRECISION_AUTO     (4) FP32 until the plan is tuned (aclfftExtTune, or ACLFFT_TUNE=1 at
 *                                     the first execution); the tuner then keeps the fastest of FP32
 *                                     and the MFMA modes whose result is within rel-L2 1e-6 of the
 *                                     FP32 result.  Plans without an MFMA mode: NOT_SUPPORTED. */
#define ACLFFT_EXT_PRECISION_AUTO 4
ACLFFT_API aclfftResult aclfftExtSetPrecision(aclfftHandle plan, int precision);

/* Measure now, on the plan's stream, the alternatives of the plan's launch list and keep the
 * fastest (like FFTW's measured planning): for uneven plans the form (one persistent launch or one
 * launch per length), the persistent kernel's schedule, workgroups per CU and compiled class set;
 * with ACLFFT_EXT_PRECISION_AUTO the operand precision; for every plan each step's compiled kernel
 * variants.  Synchronous; allocates temporary buffers of the plan's input and output size; the
 * choices are stored in the tuning cache ($ACLFFT_TUNE_CACHE, default
 * $HOME/.cache/ops_fft_amd/tune.txt), from which later plans of the same shape take them on their
 * first execution.  No counterpart in the reference, whose kernel choice is a static table
 * (src/fft1_d/arch32/fft1_d_exec_api.cpp:29-40, src/common/tiling/fft_tile_common.h:55-102). */
ACLFFT_API aclfftResult aclfftExtTune(aclfftHandle plan);

/* One plan for many C2C transforms of uneven lengths: group g holds batches[g] transforms of
 * length lengths[g], stored contiguously, groups one after the other (group g starts at the
 * complex64 offset sum_{i<g} lengths[i]*batches[i]).  Execute with aclfftExecC2C on the whole
 * buffer.  Lengths with a compiled class (60, 128, 243, 256, 512, 1000, 1024, 2048, 4096, 6561)
 * run in one persistent-kernel launch that drains a work queue of tiles; other lengths run as
 * extra launches.  Executions on device pointers replay a hipGraph captured per (in, out,
 * direction).  Errors: plan/lengths/batches null -> INVALID_VALUE; count, a length or a batch out of
 * range (length in [1, 2^27], batch >= 1) -> INVALID_SIZE; type other than C2C -> NOT_SUPPORTED.
 * No counterpart in the reference (its plans hold one length). */
ACLFFT_API aclfftResult aclfftExtPlanManyUneven(aclfftHandle* plan, int count, const int* lengths, const int* batches,
                                               aclfftType type);

/* Capture the plan's launch sequence in a hipGraph per (in, out, direction) and replay it on
 * later executions with device pointers (on = 1), or launch kernel by kernel (on = 0). */
ACLFFT_API aclfftResult aclfftExtSetGraphMode(aclfftHandle plan, int on);

/* Multiply every output of the plan by `scale` (normalisation: 1/n, 1/sqrt(n), ...), fused into
 * the last launch.  Default 1 (the reference's unnormalised convention).  Rebuilds the plan. */
ACLFFT_API aclfftResult aclfftExtSetScale(aclfftHandle plan, float scale);

/* Hand-written adjoints of the real transforms (for reverse-mode differentiation).  For an R2C
 * plan of length n: grad_x[m] = Re sum_{k=0}^{n/2} g[k] e^{+2 pi i km/n}, computed by the C2R
 * kernels with the bins strictly between DC and Nyquist weighted by 1/2.  For a C2R plan:
 * grad_X = R2C(g) with the same bins weighted by 2.  The plan's scale applies.  Same pointer rules
 * and errors as aclfftExecC2R / aclfftExecR2C; a plan of the other type -> INVALID_TYPE.  The
 * same entry points serve D2Z / Z2D plans (then the buffers hold double / complex128 values).  The
 * adjoint of a C2C transform is the C2C transform in the opposite direction. */
ACLFFT_API aclfftResult aclfftExtExecR2CAdjoint(aclfftHandle plan, aclfftComplex* grad_out, aclfftReal* grad_in);
ACLFFT_API aclfftResult aclfftExtExecC2RAdjoint(aclfftHandle plan, aclfftReal* grad_out, aclfftComplex* grad_in);

/* aclfftExecC2C that allocates nothing, for callers that cannot allocate device memory (code under
 * hipGraph stream capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in
 * the reference.  Instead of the plan's own device tables and workspace it uses the library's
 * device scratch area: OFX_SCRATCH_BYTES (64 MiB) per GPU, a device array held from the library's
 * load until the process ends; the twiddle tables are regenerated into it on every call (a few small
 * kernels), the four-step workspace follows them.  Device pointers only (host pointers ->
 * INVALID_VALUE); a plan whose tables plus workspace exceed the area -> INVALID_SIZE (with a
 * message on stderr); uneven plans -> NOT_SUPPORTED.  Asynchronous on the plan's stream.  Every call
 * on one GPU uses the same area, so calls that overlap on one GPU (several streams, several host
 * threads) race: serialise them.  The result is the plan's aclfftExecC2C result. */
ACLFFT_API aclfftResult aclfftExtExecC2CNoAlloc(aclfftHandle plan, aclfftComplex* idata, aclfftComplex* odata,
                                                int direction);

#ifdef __cplusplus
}
#endif

#endif /* OPS_FFT_EXT_H */
