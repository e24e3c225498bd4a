// j3d_resident.h -- the resident-tile persistent 3D 7-point path (j3d_resident.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "j3d_common.h"

namespace perks {

// Variant number (opts.tune[4]) of the resident-tile form of the 7-point
// persistent method.
constexpr int kResidentVariant = 5;

// The resident kernel (FP64/FP32 storage) and its workgroup size; nullptr otherwise.
void* j3d_resident_fn(int dtype, int bc, bool has_div, int* block);
// Grid for a domain: PERKS_ERR_UNSUPPORTED_SIZE when its blocks need more waves
// than cu_count x occupancy workgroups hold.
int plan_resident(int dtype, int nz, int ny, int nx, const perks_device_caps* caps, int occupancy,
                  perks_plan_info* out);
// Largest height whose ny x nx domain the resident kernel holds.
int resident_height(int dtype, int ny, int nx, const perks_device_caps* caps, int occupancy);
int resident_occupancy(int dtype, int bc, bool has_div, int* occ);
int run_resident(int dtype, void* d_a, void* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
                 const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep);

}  // namespace perks
