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

#ifdef __cplusplus
}
#endif

#endif /* OPS_FFT_EXT_H */
