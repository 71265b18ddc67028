// fft_utils_api.cpp - aclfftGetErrorString and the placeholder entry points.
//
// The error strings are the reference's (lib/fft_utils_api.cpp:19-56) so that callers comparing
// or logging them see no difference.  The workspace, stride and version queries are placeholders
// returning ACLFFT_NOT_IMPLEMENTED, as in the reference (lib/fft_utils_api.cpp:66-109).
#include "cann_ops_fft.h"

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

ACLFFT_API aclfftResult aclfftGetSize1d(aclfftHandle plan, int nx, aclfftType type, int batch, size_t* workSize) {
    (void)plan; (void)nx; (void)type; (void)batch; (void)workSize;
    return ACLFFT_NOT_IMPLEMENTED;
}
ACLFFT_API aclfftResult aclfftGetSize2d(aclfftHandle plan, int nx, int ny, aclfftType type, size_t* workSize) {
    (void)plan; (void)nx; (void)ny; (void)type; (void)workSize;
    return ACLFFT_NOT_IMPLEMENTED;
}
ACLFFT_API aclfftResult aclfftGetSize3d(aclfftHandle plan, int nx, int ny, int nz, aclfftType type,
                                        size_t* workSize) {
    (void)plan; (void)nx; (void)ny; (void)nz; (void)type; (void)workSize;
    return ACLFFT_NOT_IMPLEMENTED;
}
ACLFFT_API aclfftResult aclfftGetSize(aclfftHandle plan, size_t* workSize) {
    (void)plan; (void)workSize;
    return ACLFFT_NOT_IMPLEMENTED;
}
ACLFFT_API aclfftResult aclfftSetAutoAllocation(aclfftHandle plan, int autoAllocate) {
    (void)plan; (void)autoAllocate;
    return ACLFFT_NOT_IMPLEMENTED;
}
ACLFFT_API aclfftResult aclfftSetWorkArea(aclfftHandle plan, void* workArea) {
    (void)plan; (void)workArea;
    return ACLFFT_NOT_IMPLEMENTED;
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
