// fft_utils_api.cpp - aclfftGetErrorString and the placeholder entry points.
//
// The error strings are the reference's (lib/fft_utils_api.cpp:19-56) so that callers comparing
// or logging them see no difference.  The workspace API is implemented (below); the stride and
// version queries are placeholders returning ACLFFT_NOT_IMPLEMENTED, as in the reference
// (lib/fft_utils_api.cpp:96-109).
#include "cann_ops_fft.h"
#include "executor.h"
#include "fft_handle_impl.h"
#include "fft_validate.h"

extern "C" {

ACLFFT_API const char* aclfftGetErrorString(aclfftResult result) {
    switch (result) {
        case ACLFFT_SUCCESS: return "Success";
        case ACLFFT_INVALID_PLAN: return "Invalid plan handle";
        case ACLFFT_ALLOC_FAILED: return "Memory allocation failed";
        case ACLFFT_INVALID_TYPE: return "Invalid data type";
        case ACLFFT_INVALID_VALUE: return "Invalid parameter value";
        case ACLFFT_INTERNAL_ERROR: return "Internal error";
        case ACLFFT_EXEC_FAILED: return "Execution failed";
        case ACLFFT_SETUP_FAILED: return "Setup failed";
        case ACLFFT_INVALID_SIZE: return "Invalid size parameter";
        case ACLFFT_UNALIGNED_DATA: return "Data not aligned";
        case ACLFFT_INCOMPLETE_PARAMETER_LIST: return "Incomplete parameter list";
        case ACLFFT_INVALID_DEVICE: return "Invalid device";
        case ACLFFT_PARSE_ERROR: return "Parse error";
        case ACLFFT_NO_WORKSPACE: return "No workspace available";
        case ACLFFT_NOT_IMPLEMENTED: return "Feature not implemented";
        case ACLFFT_NOT_SUPPORTED: return "Feature not supported";
        default: return "Unknown error";
    }
}

// ---------------------------------------------------------------------------------------------
// Workspace API.  The reference declares these (include/cann_ops_fft.h:258-336) and ships them as
// NOT_IMPLEMENTED placeholders (lib/fft_utils_api.cpp:66-94); its roadmap
// (docs/implementation.md:193-204) asks for automatic, caller-managed and shared workspaces.  Here:
//   aclfftGetSize           the built plan's workspace (bytes; 0 when it needs none)
//   aclfftGetSize1d / 2d    the workspace a plan of these parameters would need (host-only build of
//                           the launch list; same validation and codes as MakePlan1d / 2d; the
//                           2D query is for one transform, as the call has no batch argument)
//   aclfftGetSize3d         the same for 3D plans (the reference's is a placeholder, kept under
//                           ACLFFT_STRICT_REFERENCE_DOMAIN=1)
//   aclfftSetAutoAllocation 0: the plan never allocates its workspace; the caller provides it
//   aclfftSetWorkArea       device memory of at least aclfftGetSize bytes, used from then on
//                           (several plans may share one area when they run on one stream)
// An execution that needs workspace while automatic allocation is off and no area is set returns
// ACLFFT_NO_WORKSPACE.  Host-pointer executions still stage through plan-owned buffers.
// ---------------------------------------------------------------------------------------------

namespace {

aclfftResult size_of(const ofx::PlanDesc& d, size_t* workSize) {
    ofx::Executor* e = nullptr;
    aclfftResult r = ofx::Executor::build(d, &e);
    if (r == ACLFFT_NOT_IMPLEMENTED) {
        *workSize = 0;  // valid plan without an execution path: nothing would be allocated
        return ACLFFT_SUCCESS;
    }
    if (r != ACLFFT_SUCCESS) return r;
    *workSize = e->work_bytes();
    delete e;
    return ACLFFT_SUCCESS;
}

}  // namespace

ACLFFT_API aclfftResult aclfftGetSize1d(aclfftHandle plan, int nx, aclfftType type, int batch, size_t* workSize) {
    if (plan == nullptr) return ACLFFT_INVALID_PLAN;
    if (workSize == nullptr) return ACLFFT_INVALID_VALUE;
    aclfftResult r = ofx::check_1d_args(nx, type, batch, ACLFFT_HORIZONTAL, "aclfftGetSize1d");
    if (r != ACLFFT_SUCCESS) return r;
    ofx::PlanDesc d;
    d.rank = 1;
    d.n0 = nx;
    d.batch = batch;
    d.type = type;
    d.precision = plan->precision;
    return size_of(d, workSize);
}

ACLFFT_API aclfftResult aclfftGetSize2d(aclfftHandle plan, int nx, int ny, aclfftType type, size_t* workSize) {
    if (plan == nullptr) return ACLFFT_INVALID_PLAN;
    if (workSize == nullptr) return ACLFFT_INVALID_VALUE;
    if (nx <= 0 || ny <= 0) return ACLFFT_INVALID_SIZE;
    if (!ofx::is_known_type(type)) return ACLFFT_INVALID_TYPE;
    if (ofx::check_2d_type(type, "aclfftGetSize2d") != ACLFFT_SUCCESS) return ACLFFT_NOT_SUPPORTED;
    aclfftResult r = ofx::check_2d_domain(nx, ny, "aclfftGetSize2d");
    if (r != ACLFFT_SUCCESS) return r;
    ofx::PlanDesc d;
    d.rank = 2;
    d.n0 = nx;
    d.n1 = ny;
    d.batch = 1;
    d.type = type;
    return size_of(d, workSize);
}

// The reference's is a stub (lib/fft_utils_api.cpp:76-79), kept in strict mode; otherwise the
// checks of aclfftGetSize2d with the 3D gates of aclfftMakePlan3d.
ACLFFT_API aclfftResult aclfftGetSize3d(aclfftHandle plan, int nx, int ny, int nz, aclfftType type,
                                        size_t* workSize) {
    if (ofx::strict_reference_domain()) return ACLFFT_NOT_IMPLEMENTED;
    if (plan == nullptr) return ACLFFT_INVALID_PLAN;
    if (workSize == nullptr) return ACLFFT_INVALID_VALUE;
    if (nx <= 0 || ny <= 0 || nz <= 0) return ACLFFT_INVALID_SIZE;
    if (!ofx::is_known_type(type)) return ACLFFT_INVALID_TYPE;
    aclfftResult r = ofx::check_3d_type_domain(nx, ny, nz, type, "aclfftGetSize3d");
    if (r != ACLFFT_SUCCESS) return r;
    ofx::PlanDesc d;
    d.rank = 3;
    d.n0 = nx;
    d.n1 = ny;
    d.n2 = nz;
    d.batch = 1;
    d.type = type;
    return size_of(d, workSize);
}

ACLFFT_API aclfftResult aclfftGetSize(aclfftHandle plan, size_t* workSize) {
    if (plan == nullptr || !plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (workSize == nullptr) return ACLFFT_INVALID_VALUE;
    *workSize = plan->exec ? plan->exec->work_bytes() : 0;
    return ACLFFT_SUCCESS;
}

ACLFFT_API aclfftResult aclfftSetAutoAllocation(aclfftHandle plan, int autoAllocate) {
    if (plan == nullptr) return ACLFFT_INVALID_PLAN;
    plan->auto_allocate = autoAllocate != 0 ? 1 : 0;
    if (plan->exec) plan->exec->set_auto_allocation(autoAllocate != 0);
    return ACLFFT_SUCCESS;
}

ACLFFT_API aclfftResult aclfftSetWorkArea(aclfftHandle plan, void* workArea) {
    if (plan == nullptr || !plan->is_initialized) return ACLFFT_INVALID_PLAN;
    if (workArea == nullptr) return ACLFFT_INVALID_VALUE;
    if (plan->exec == nullptr) return ACLFFT_SUCCESS;  // plan without an execution path: nothing to bind
    if (plan->exec->work_bytes() == 0) return ACLFFT_SUCCESS;
    return plan->exec->set_work_area(workArea);
}

ACLFFT_API aclfftResult aclfftSetStride(aclfftHandle plan, int32_t stride) {
    (void)plan; (void)stride;
    return ACLFFT_NOT_IMPLEMENTED;
}
ACLFFT_API aclfftResult aclfftGetVersion(int* version) {
    (void)version;
    return ACLFFT_NOT_IMPLEMENTED;
}
ACLFFT_API aclfftResult aclfftGetProperty(aclfftLibraryPropertyType type, int* value) {
    (void)type; (void)value;
    return ACLFFT_NOT_IMPLEMENTED;
}

}  // extern "C"
