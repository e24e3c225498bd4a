// j3d_kernels.h -- host-side launchers of the 3D Jacobi kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "j3d_common.h"

namespace perks {

struct PlainGeom {
    int grid;
    int block;
    int zchunk;
};


// One time step of the 3D star r1 (7-point) stencil: out = S(in).
// variant bit 0: nontemporal stores.
template <typename T>
int launch_j3d_star1_plain(const T* in, T* out, int nz, int ny, int nx, const KCoefs& cf, int bc, int zchunk,
                           hipStream_t stream, PlainGeom* geom, int variant);

// The same step on FP16 (fmt PERKS_F16) or BF16 (PERKS_BF16) storage with FP32
// arithmetic; nx must be even.
// sr: stochastic rounding with the time step's hash key (storage16.h).
int launch_j3d_star1_plain16(int fmt, const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const KCoefs& cf,
                             int bc, int zchunk, bool sr, uint32_t key, hipStream_t stream, PlainGeom* geom);

// The kernels the launchers above run (256-thread workgroups), for attribute
// and occupancy queries.
void* j3d_star1_plain_fn(int dtype, int bc, bool has_div, bool nontemporal);
void* j3d_star1_plain16_fn(int fmt, int bc, bool has_div, bool sr);

}  // namespace perks
