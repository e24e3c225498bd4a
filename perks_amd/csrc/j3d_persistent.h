// j3d_persistent.h -- the persistent (one cooperative launch) 3D Jacobi path.
#pragma once

#include <hip/hip_runtime.h>

#include "j3d_common.h"

namespace perks {

template <typename T>
int run_persistent(T* d_a, T* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
                   const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep);

int release_persistent_scratch();

}  // namespace perks
