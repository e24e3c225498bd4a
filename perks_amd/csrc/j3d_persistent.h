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
// The persistent kernel of tile variant `var` (fused: the contracted update),
// and its workgroup size in *block, for attribute and occupancy queries;
// nullptr when not built.
void* j3d_persistent_fn(int dtype, int bc, bool has_div, int var, bool sr, bool fused, int* block);

}  // namespace perks
