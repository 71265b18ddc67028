// fft_validate.h - argument validation shared by the plan entry points.
//
// The checks, their order and their result codes follow the reference plan layer so that every
// rejection its negative tests assert produces the same aclfftResult:
//   lib/fft_plan_api.cpp:47-69     (aclfftPlan1d)      lib/fft_plan_init_api.cpp:118-176 (MakePlan1d)
//   lib/fft_plan_api.cpp:90-105    (aclfftPlan2d)      lib/fft_plan_init_api.cpp:189-236 (MakePlan2d)
//   lib/fft_plan_validate.h:20-72  (1D size domain, 2D size domain)
//   lib/fft_error.h:75-87          (type whitelist)
// Rejections are reported on stderr with the "[ops-fft] <entry>:" prefix the reference uses.
#ifndef OFX_FFT_VALIDATE_H
#define OFX_FFT_VALIDATE_H

#include <cstdio>
#include <cstdlib>
#include "cann_ops_fft.h"

namespace ofx {

constexpr int kMaxHorizontalN = 1 << 27;  // horizontal 1D length bound
constexpr int kMinVerticalN = 256;        // vertical 1D lengths: powers of two in [2^8, 2^18]
constexpr int kMaxVerticalN = 1 << 18;

inline bool is_known_type(aclfftType t) {
    return t == ACLFFT_C2C || t == ACLFFT_R2C || t == ACLFFT_C2R || t == ACLFFT_Z2Z || t == ACLFFT_D2Z ||
           t == ACLFFT_Z2D;
}
inline bool is_single_precision_type(aclfftType t) {
    return t == ACLFFT_C2C || t == ACLFFT_R2C || t == ACLFFT_C2R;
}

inline void log_reject(const char* entry, const char* what) { std::fprintf(stderr, "[ops-fft] %s: %s\n", entry, what); }

// When ACLFFT_STRICT_REFERENCE_DOMAIN=1 the plan entry points keep the reference's domain exactly:
// 2D plans limited to {32,64,128}^2 (lib/fft_plan_validate.h:52-72) and double precision refused
// (lib/fft_plan_api.cpp:58-62).  By default any 2D nx, ny in [1, 2^27] and 1D double precision
// are accepted.
inline bool strict_reference_domain() {
    const char* v = std::getenv("ACLFFT_STRICT_REFERENCE_DOMAIN");
    return v != nullptr && v[0] == '1';
}

// 1D length domain by layout (reference lib/fft_plan_validate.h:30-47).
inline aclfftResult check_1d_domain(int dimType, int nx, const char* entry) {
    char msg[160];
    if (dimType == ACLFFT_HORIZONTAL) {
        if (nx > kMaxHorizontalN) {
            std::snprintf(msg, sizeof msg, "nx=%d exceeds 2^27 for horizontal FFT", nx);
            log_reject(entry, msg);
            return ACLFFT_INVALID_SIZE;
        }
        return ACLFFT_SUCCESS;
    }
    const bool pow2 = nx > 0 && (nx & (nx - 1)) == 0;
    if (!pow2 || nx < kMinVerticalN || nx > kMaxVerticalN) {
        std::snprintf(msg, sizeof msg, "nx=%d out of [256, 262144] or not a power of 2 for vertical FFT", nx);
        log_reject(entry, msg);
        return ACLFFT_INVALID_SIZE;
    }
    return ACLFFT_SUCCESS;
}

// The common 1D argument chain: sizes, layout, type whitelist, single-precision gate, domain.
inline aclfftResult check_1d_args(int nx, aclfftType type, int batch, int dimType, const char* entry) {
    if (nx <= 0) return ACLFFT_INVALID_SIZE;
    if (batch <= 0) return ACLFFT_INVALID_SIZE;
    if (dimType != ACLFFT_HORIZONTAL && dimType != ACLFFT_VERTICAL) return ACLFFT_INVALID_VALUE;
    if (!is_known_type(type)) return ACLFFT_INVALID_TYPE;
    if (!is_single_precision_type(type) && strict_reference_domain()) {
        // the reference refuses double precision at plan time (lib/fft_plan_api.cpp:58-62); by
        // default this library accepts it (1D Z2Z / D2Z / Z2D)
        char msg[96];
        std::snprintf(msg, sizeof msg, "type=%d has no 1D execution implementation yet", (int)type);
        log_reject(entry, msg);
        return ACLFFT_NOT_SUPPORTED;
    }
    return check_1d_domain(dimType, nx, entry);
}

// 2D type gate: the reference implements 2D C2C only (lib/fft_plan_api.cpp:96-100,
// lib/fft_plan_init_api.cpp:200-204); here 2D Z2Z runs too, except in strict mode.
inline aclfftResult check_2d_type(aclfftType type, const char* entry) {
    if (type == ACLFFT_C2C || (type == ACLFFT_Z2Z && !strict_reference_domain())) return ACLFFT_SUCCESS;
    char msg[96];
    std::snprintf(msg, sizeof msg, "type=%d has no 2D execution implementation yet", (int)type);
    log_reject(entry, msg);
    return ACLFFT_NOT_SUPPORTED;
}

inline aclfftResult check_2d_domain(int nx, int ny, const char* entry) {
    char msg[160];
    if (strict_reference_domain()) {
        auto dd = [](int s) { return s == 32 || s == 64 || s == 128; };
        if (!dd(nx) || !dd(ny)) {
            std::snprintf(msg, sizeof msg, "size=%dx%d not supported, only {32,64,128}x{32,64,128} is supported", nx,
                          ny);
            log_reject(entry, msg);
            return ACLFFT_NOT_SUPPORTED;
        }
        return ACLFFT_SUCCESS;
    }
    if (nx > kMaxHorizontalN || ny > kMaxHorizontalN) {
        std::snprintf(msg, sizeof msg, "size=%dx%d exceeds 2^27 in one dimension", nx, ny);
        log_reject(entry, msg);
        return ACLFFT_NOT_SUPPORTED;
    }
    return ACLFFT_SUCCESS;
}

// 3D: the reference's aclfftMakePlan3d / aclfftGetSize3d are stubs returning NOT_IMPLEMENTED
// (lib/fft_plan_init_api.cpp:250-254, lib/fft_utils_api.cpp:76-79; kept in strict mode).  Here 3D
// C2C and Z2Z run, with the 2D gates per axis: NOT_SUPPORTED for another type or an axis beyond
// 2^27.
inline aclfftResult check_3d_type_domain(int nx, int ny, int nz, aclfftType type, const char* entry) {
    char msg[160];
    if (type != ACLFFT_C2C && type != ACLFFT_Z2Z) {
        std::snprintf(msg, sizeof msg, "type=%d has no 3D execution implementation yet", (int)type);
        log_reject(entry, msg);
        return ACLFFT_NOT_SUPPORTED;
    }
    if (nx > kMaxHorizontalN || ny > kMaxHorizontalN || nz > kMaxHorizontalN) {
        std::snprintf(msg, sizeof msg, "size=%dx%dx%d exceeds 2^27 in one dimension", nx, ny, nz);
        log_reject(entry, msg);
        return ACLFFT_NOT_SUPPORTED;
    }
    return ACLFFT_SUCCESS;
}

}  // namespace ofx

#endif
