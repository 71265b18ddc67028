/*
 * cann_ops_fft.h - public C API of the MI355X (gfx950) FFT library.
 *
 * Drop-in replacement for the aclfft API of ops-fft (include/cann_ops_fft.h:19-477 of the
 * reference).  Every enum value, type, struct layout and prototype below is identical to the
 * reference header, with one deliberate change: the stream handle of aclfftSetStream is a HIP
 * stream (hipStream_t) instead of an ACL stream (aclrtStream).  Both are opaque pointers, so the
 * ABI (calling convention and argument sizes) is unchanged.
 *
 * Pointer contract (reference: docs/zh/API_Reference/FFT_Common_Interface.md:5, host pointers,
 * synchronous):
 *   - idata/odata that are host pointers are staged through plan-owned device buffers and the
 *     call returns after the result has been copied back (the reference's semantics);
 *   - idata/odata that are device pointers (hipMalloc'd) run asynchronously on the plan's stream
 *     with no staging.  Both pointers of one call must be of the same kind.
 *
 * Library additions that have no counterpart in the reference are declared in
 * ops_fft_ext.h; this header stays limited to the reference's 25 entry points.
 */
#ifndef CANN_OPS_FFT_H
#define CANN_OPS_FFT_H

#include <stdint.h>
#include <stddef.h>

/* hipStream_t, declared the way hip_runtime_api.h declares it so that C callers do not need the
 * HIP headers (a repeated identical typedef is valid C11 and C++). */
struct ihipStream_t;
typedef struct ihipStream_t* hipStream_t;

#ifndef ACLFFT_NO_ACLRT_STREAM_ALIAS
/* Source compatibility for callers written against the reference (aclrtStream came from
 * acl/acl.h there). */
typedef hipStream_t aclrtStream;
#endif

#if defined(_WIN32) || defined(__CYGWIN__)
#define ACLFFT_API __declspec(dllexport)
#else
#define ACLFFT_API __attribute__((visibility("default")))
#endif

/* Result codes: values identical to the reference (include/cann_ops_fft.h:32-66; 15 unused). */
typedef enum aclfftResult_t {
    ACLFFT_SUCCESS = 0,
    ACLFFT_INVALID_PLAN = 1,
    ACLFFT_ALLOC_FAILED = 2,
    ACLFFT_INVALID_TYPE = 3,
    ACLFFT_INVALID_VALUE = 4,
    ACLFFT_INTERNAL_ERROR = 5,
    ACLFFT_EXEC_FAILED = 6,
    ACLFFT_SETUP_FAILED = 7,
    ACLFFT_INVALID_SIZE = 8,
    ACLFFT_UNALIGNED_DATA = 9,
    ACLFFT_INCOMPLETE_PARAMETER_LIST = 10,
    ACLFFT_INVALID_DEVICE = 11,
    ACLFFT_PARSE_ERROR = 12,
    ACLFFT_NO_WORKSPACE = 13,
    ACLFFT_NOT_IMPLEMENTED = 14,
    ACLFFT_NOT_SUPPORTED = 16
} aclfftResult;

/* Transform types (reference include/cann_ops_fft.h:80-94). */
typedef enum aclfftType_t {
    ACLFFT_R2C = 0x2a,
    ACLFFT_C2R = 0x2c,
    ACLFFT_C2C = 0x29,
    ACLFFT_D2Z = 0x6a,
    ACLFFT_Z2D = 0x6c,
    ACLFFT_Z2Z = 0x69
} aclfftType;

/* reference include/cann_ops_fft.h:96-101 */
typedef enum aclfftLibraryPropertyType_t {
    ACLFFT_MAJOR_VERSION,
    ACLFFT_MINOR_VERSION,
    ACLFFT_PATCH_LEVEL
} aclfftLibraryPropertyType;

/* reference include/cann_ops_fft.h:105-111 */
#define ACLFFT_FORWARD (-1)
#define ACLFFT_BACKWARD 1

#define ACLFFT_HORIZONTAL 0
#define ACLFFT_VERTICAL 1

typedef struct aclfftHandle_t* aclfftHandle;

/* Interleaved complex64: x = real, y = imaginary (reference include/cann_ops_fft.h:117-120). */
typedef struct {
    float x;
    float y;
} aclfftComplex;

typedef struct {
    double x;
    double y;
} aclfftDoubleComplex;

typedef float aclfftReal;
typedef double aclfftDoubleReal;

#ifdef __cplusplus
extern "C" {
#endif

/* Create + initialise a 1D plan.  nx: transform length (HORIZONTAL: 1..2^27; VERTICAL: power of
 * two in [256, 2^18]); batch: number of transforms; dimType: ACLFFT_HORIZONTAL or
 * ACLFFT_VERTICAL.  Replaces reference include/cann_ops_fft.h:152-156 (lib/fft_plan_api.cpp:47-85). */
ACLFFT_API aclfftResult aclfftPlan1d(aclfftHandle* plan, int nx, aclfftType type, int batch, int dimType);

/* Create + initialise a 2D C2C (or Z2Z) plan, row-major [batch][nx][ny] (ny fastest).  Argument
 * order is (plan, batch, nx, ny, type) as in the reference include/cann_ops_fft.h:174
 * (lib/fft_plan_api.cpp:90-122). */
ACLFFT_API aclfftResult aclfftPlan2d(aclfftHandle* plan, int batch, int nx, int ny, aclfftType type);

/* Create + initialise a 3D C2C / Z2Z plan of one [nx][ny][nz] volume (nz fastest).  Replaces the
 * reference's placeholder (include/cann_ops_fft.h:192-193, lib/fft_plan_api.cpp:127-150), whose
 * argument checks come first here too; other types: ACLFFT_NOT_SUPPORTED.  With
 * ACLFFT_STRICT_REFERENCE_DOMAIN=1 it returns ACLFFT_NOT_IMPLEMENTED like the reference. */
ACLFFT_API aclfftResult aclfftPlan3d(aclfftHandle* plan, int nx, int ny, int nz, aclfftType type);

/* Allocate an empty plan (reference include/cann_ops_fft.h:199, lib/fft_plan_api.cpp:24-36). */
ACLFFT_API aclfftResult aclfftCreate(aclfftHandle* plan);

/* Initialise a created plan.  *workSize (if non-null) receives the bytes of device workspace the
 * plan holds (the reference always reports 0: lib/fft_plan_init_api.cpp:167-170).
 * Replaces reference include/cann_ops_fft.h:213-218 (lib/fft_plan_init_api.cpp:118-176). */
ACLFFT_API aclfftResult aclfftMakePlan1d(aclfftHandle plan, int nx, aclfftType type, int batch, int dimType,
                                         size_t* workSize);

/* Replaces reference include/cann_ops_fft.h:238-239 (lib/fft_plan_init_api.cpp:189-236).  Unlike
 * the reference, any nx, ny in the 1D length domain is accepted (the reference only has
 * {32,64,128}^2). */
ACLFFT_API aclfftResult aclfftMakePlan2d(aclfftHandle plan, int batch, int nx, int ny, aclfftType type,
                                         size_t* workSize);

/* Replaces the reference's placeholder (include/cann_ops_fft.h:259-260,
 * lib/fft_plan_init_api.cpp:250-254): the 3D plan of aclfftPlan3d, NOT_IMPLEMENTED in strict mode. */
ACLFFT_API aclfftResult aclfftMakePlan3d(aclfftHandle plan, int nx, int ny, int nz, aclfftType type,
                                         size_t* workSize);

/* Workspace API (include/cann_ops_fft.h:273-336; placeholders in the reference,
 * lib/fft_utils_api.cpp:66-94, implemented here): GetSize1d/2d/3d report the workspace a plan of
 * those arguments would hold (host only), GetSize the built plan's; SetAutoAllocation(plan, 0) +
 * SetWorkArea(plan, devptr) run the plan on the caller's device area. */
ACLFFT_API aclfftResult aclfftGetSize1d(aclfftHandle plan, int nx, aclfftType type, int batch, size_t* workSize);
ACLFFT_API aclfftResult aclfftGetSize2d(aclfftHandle plan, int nx, int ny, aclfftType type, size_t* workSize);
ACLFFT_API aclfftResult aclfftGetSize3d(aclfftHandle plan, int nx, int ny, int nz, aclfftType type,
                                        size_t* workSize);
ACLFFT_API aclfftResult aclfftGetSize(aclfftHandle plan, size_t* workSize);
ACLFFT_API aclfftResult aclfftSetAutoAllocation(aclfftHandle plan, int autoAllocate);
ACLFFT_API aclfftResult aclfftSetWorkArea(aclfftHandle plan, void* workArea);

/* Complex-to-complex FP32 transform, unnormalised in both directions
 * (reference include/cann_ops_fft.h:348-351, lib/fft_exec_api.cpp:32-51). */
ACLFFT_API aclfftResult aclfftExecC2C(aclfftHandle plan, aclfftComplex* idata, aclfftComplex* odata,
                                      int direction);

/* Real-to-complex forward: [batch][nx] float -> [batch][nx/2+1] complex64
 * (reference include/cann_ops_fft.h:363-365, lib/fft_exec_api.cpp:53-68). */
ACLFFT_API aclfftResult aclfftExecR2C(aclfftHandle plan, aclfftReal* idata, aclfftComplex* odata);

/* Complex-to-real inverse, unnormalised: [batch][nx/2+1] complex64 -> [batch][nx] float.  The
 * imaginary parts of the DC and Nyquist bins are ignored
 * (reference include/cann_ops_fft.h:376-378, lib/fft_exec_api.cpp:70-85). */
ACLFFT_API aclfftResult aclfftExecC2R(aclfftHandle plan, aclfftComplex* idata, aclfftReal* odata);

/* Double precision: complex128 / float64 counterparts of ExecC2C / ExecR2C / ExecC2R on plans of
 * type Z2Z / D2Z / Z2D.  The reference ships these as NOT_IMPLEMENTED placeholders and refuses such
 * plans (lib/fft_exec_api.cpp:87-116, lib/fft_plan_api.cpp:58-62); here they run on the complex128
 * kernels (ACLFFT_STRICT_REFERENCE_DOMAIN=1 restores the refusal). */
ACLFFT_API aclfftResult aclfftExecZ2Z(aclfftHandle plan, aclfftDoubleComplex* idata, aclfftDoubleComplex* odata,
                                      int direction);
ACLFFT_API aclfftResult aclfftExecD2Z(aclfftHandle plan, aclfftDoubleReal* idata, aclfftDoubleComplex* odata);
ACLFFT_API aclfftResult aclfftExecZ2D(aclfftHandle plan, aclfftDoubleComplex* idata, aclfftDoubleReal* odata);

/* Associate a HIP stream with the plan (reference include/cann_ops_fft.h:429,
 * lib/fft_stream_api.cpp:22-32; the reference takes an aclrtStream). */
ACLFFT_API aclfftResult aclfftSetStream(aclfftHandle plan, hipStream_t stream);

ACLFFT_API aclfftResult aclfftSetStride(aclfftHandle plan, int32_t stride);

/* Destroy a plan and release its device buffers (reference include/cann_ops_fft.h:445,
 * lib/fft_plan_destroy_api.cpp:24-40). */
ACLFFT_API aclfftResult aclfftDestroy(aclfftHandle plan);

ACLFFT_API aclfftResult aclfftGetVersion(int* version);
ACLFFT_API aclfftResult aclfftGetProperty(aclfftLibraryPropertyType type, int* value);

/* Error strings identical to reference lib/fft_utils_api.cpp:19-56. */
ACLFFT_API const char* aclfftGetErrorString(aclfftResult result);

#ifdef __cplusplus
}
#endif

#endif /* CANN_OPS_FFT_H */
