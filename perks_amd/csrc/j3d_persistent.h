// j3d_persistent.h -- the persistent (one cooperative launch) 3D Jacobi path.
#pragma once

#include <hip/hip_runtime.h>

#include "j3d_common.h"

namespace perks {

template <typename T>
int run_persistent(T* d_a, T* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
                   const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep);

// FP16 (fmt PERKS_F16) or BF16 (PERKS_BF16) storage, FP32 arithmetic.
int run_persistent16(int fmt, uint16_t* d_a, uint16_t* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
                     const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep);

int release_persistent_scratch();

// Number of tile variants (opts.tune[4]) of the persistent 7-point kernel.
int j3d_persistent_variants();
// The form of the 7-point update a persistent run uses (j3dp::kUpd*): the
// contracted one when opts.tune[6] = 1; the shared-product one (the gold's
// bits) when the six neighbour weights are one positive value in the
// arithmetic type (FP64, or FP32 for FP32/FP16/BF16 storage); the gold's
// otherwise.
int j3d_persistent_update(int dtype, const perks_stencil_opts* o, const KCoefs& cf);
// The persistent kernel of tile variant `var` and update form `up`, and its
// workgroup size in *block, for attribute and occupancy queries; nullptr when
// not built.
void* j3d_persistent_fn(int dtype, int bc, bool has_div, int var, bool sr, int up, int* block);

}  // namespace perks
