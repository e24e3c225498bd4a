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

}  // namespace perks
