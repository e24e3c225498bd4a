// cg.h -- persistent conjugate gradient (host side).
#pragma once

#include "perks_amd.h"

namespace perks {

// fmt 0 = CSR (csr), 1 = DIA (dia).
int cg_solve(int fmt, const perks_csr* csr, const perks_dia* dia, const double* d_b, double* d_x,
             const perks_cg_opts* o, void* stream, perks_cg_report* rep);
int cg_solve_f32(int fmt, const perks_csr_f32* csr, const perks_dia_f32* dia, const float* d_b, float* d_x,
                 const perks_cg_opts* o, void* stream, perks_cg_report* rep);
int poisson7_csr(int nx, int ny, int nz, int32_t* rowptr, int32_t* col, double* val, void* stream);
int poisson7_dia(int nx, int ny, int nz, double* val, void* stream);
int release_cg_scratch();

}  // namespace perks
