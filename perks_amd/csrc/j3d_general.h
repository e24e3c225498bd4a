// j3d_general.h -- the shape-generic 3D Jacobi path (every 3D shape, every
// storage format; per-step and persistent methods).
#pragma once

#include <hip/hip_runtime.h>

#include "j3d_common.h"

namespace perks {

// Runs `steps` time steps on device buffers d_a/d_b of format dtype; the
// result is in d_b when rep->result_in_b.
int j3g_run(int dtype, void* d_a, void* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
            const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep);
// Host-only plan (no device call): grid, tile, LDS bytes; reserved[0] = steps
// per pass, reserved[1] = z-slabs per tile column.
int j3g_plan(int dtype, int nz, int ny, int nx, const perks_stencil_opts* o, const perks_device_caps* caps,
             int occupancy, perks_plan_info* out);
// Time steps per pass the persistent method uses for this shape and format.
int j3g_levels(int dtype, int shape, const perks_stencil_opts* o);
// The kernel o selects (opts.tune[4]: register streaming or LDS planes), its
// workgroup size and dynamic LDS bytes, for attribute and occupancy queries.
void* j3g_fn(int dtype, const perks_stencil_opts* o, int* block, size_t* lds);

}  // namespace perks
