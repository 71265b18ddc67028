// tune_cache.h - persisted autotuning choices, keyed by (type, rank, lengths, batch, step).
//
// File: $ACLFFT_TUNE_CACHE, default $HOME/.cache/ops_fft_amd/tune.txt.  One line per step:
//   <type> <rank> <n0> <n1> <n2> <batch> <vertical> <step> [u<groups hash>] <signature>
// (the hash of the (length, batch) list for uneven plans), where the signature names the kernel
// configuration (kind, n, T, TPB, staged, I/O width, radices, ":s" for the streaming variants,
// and the LDS footprint, which tells apart the column kernels' split-exchange variants and the row
// kernels' pad shifts), so a cache stays valid across rebuilds that reorder the registries.
// Row-kernel and column-kernel steps are recorded.  The cache is the only state this library
// persists.
#ifndef OFX_TUNE_CACHE_H
#define OFX_TUNE_CACHE_H

#include <string>
#include <vector>
#include "executor.h"

namespace ofx {

std::string kernel_signature(const ofx_kernel_info& k);
// The registries a plan's steps index: row kernels (KERNEL steps) and column kernels (COLUMN steps).
struct Registries {
    const ofx_kernel_info* row;
    int nrow;
    const ofx_kernel_info* col;
    int ncol;
};
// Replace each step's kernel by a cached choice when one exists among its candidates.
void apply_tune_cache(const PlanDesc& d, std::vector<Step>& steps, const Registries& reg);
// Record the steps' current choices.
void store_tune_cache(const PlanDesc& d, const std::vector<Step>& steps, const Registries& reg);
// Plan-level choices (uneven-plan knobs, automatic precision): the chosen knobs are written into
// *out (a copy of d) when the cache holds an entry for d's shape.
bool load_plan_choice(const PlanDesc& d, PlanDesc* out);
void store_plan_choice(const PlanDesc& chosen);

}  // namespace ofx

#endif
