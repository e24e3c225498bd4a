/*
 * perks_oracle.h -- CPU oracle for the Jacobi stencil and CG paths.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product library (libperks_amd.so) never
 * links or calls it.
 *
 * What it restates (reference paths relative to the PERKS source tree):
 *   - stencil/3dstencil/3d7pt/3d7pt_gold_gen.cpp:9-36   j3d_step (clamp, every cell)
 *   - stencil/3dstencil/3d7pt/3d7pt_gold_gen.cpp:38-67  j3d_gold_iterative (ping-pong parity)
 *   - stencil/3dstencil/3d7pt/3d7pt_gold.cpp:8-25       interior-only step (Dirichlet form)
 *   - stencil/3dstencil/{3d13pt,3d17pt,3d27pt,poisson}/ *_gold_gen.cpp   other 3D shapes
 *   - stencil/2dstencil/2d*pt/jacobi_gold_gen.cpp       2D shapes (a 2D grid is nz == 1)
 *   - stencil/3dstencil/common/common.hpp:24-27,105-163 rand()-based input and error metrics
 *   - conjugateGradient/cg_kernels.cuh:721-975          CG algorithm order (here in FP64,
 *     with the reduction order fixed as documented at po_dot_f64)
 *
 * A stencil is an ordered list of "groups".  Group g contributes
 *     coef_g * (x[p_0] + x[p_1] + ... + x[p_{n-1}])          (points summed left to right)
 * and the groups are accumulated left to right starting from the first group's
 * product, exactly as a C expression  c0*a + c1*b + ... would be evaluated.
 * An optional final division by `divisor` follows.  For float storage a group
 * flagged prod_f64 forms its product in double (the reference writes such
 * coefficients as double literals, e.g. `15.0*input`), and from the first such
 * group on the running sum is held in double; the division is then in double
 * too and the result is rounded to float on store.  For double storage every
 * flag is irrelevant.
 */
#ifndef PERKS_ORACLE_H
#define PERKS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PO_MAX_GROUPS 64
#define PO_MAX_POINTS 12

enum { PO_BC_CLAMP = 0, PO_BC_ZERO = 1 };

typedef struct po_group {
    double coef;
    int32_t npts;
    int32_t prod_f64;
    int32_t off[PO_MAX_POINTS][3]; /* (dz, dy, dx) */
} po_group;

typedef struct po_stencil {
    int32_t ngroups;
    int32_t boundary; /* PO_BC_CLAMP or PO_BC_ZERO */
    double divisor;   /* 0 => no division */
    po_group groups[PO_MAX_GROUPS];
} po_stencil;

/* ---- inputs ---------------------------------------------------------- */
uint64_t po_splitmix64(uint64_t seed, uint64_t index);
void po_fill_uniform_f64(double* out, size_t n, uint64_t seed, double lo, double hi);
void po_fill_uniform_f32(float* out, size_t n, uint64_t seed, double lo, double hi);
/* the reference's generator: (T)rand()/(T)(RAND_MAX-1) after srand(seed) */
void po_fill_libc_rand_f64(double* out, size_t n, unsigned seed);
void po_fill_libc_rand_f32(float* out, size_t n, unsigned seed);

/* ---- stencil --------------------------------------------------------- */
void po_stencil_step_f64(const double* in, double* out, int nz, int ny, int nx, const po_stencil* st);
void po_stencil_step_f32(const float* in, float* out, int nz, int ny, int nx, const po_stencil* st);
/* Runs `steps` steps.  The result always lands in `out` (as the reference's
 * j3d_gold_iterative guarantees); `in` is not modified.  steps == 0 copies. */
void po_stencil_iterate_f64(const double* in, double* out, int nz, int ny, int nx, const po_stencil* st, int steps);
void po_stencil_iterate_f32(const float* in, float* out, int nz, int ny, int nx, const po_stencil* st, int steps);
/* Same, with the z loop split across `threads` OpenMP threads (bitwise equal). */
void po_stencil_iterate_omp_f64(const double* in, double* out, int nz, int ny, int nx, const po_stencil* st, int steps, int threads);
void po_stencil_iterate_omp_f32(const float* in, float* out, int nz, int ny, int nx, const po_stencil* st, int steps, int threads);
/* The 3D 7-point star as the gold's own loop (3d7pt_gold_gen.cpp:9-40): one
 * cell at a time, terms e, w, n, s, t, b, c with weights k[0..6], each product
 * and sum rounded once; zero or clamp boundary; z planes split across
 * `threads` OpenMP threads.  Bitwise equal to po_stencil_iterate_f64 with
 * the same weights (the CPU baseline's all-cores figure times it). */
void po_star7_iterate_omp_f64(const double* in, double* out, int nz, int ny, int nx, const double* k, int boundary,
                              int steps, int threads);
/* Contracted form (no reference counterpart): group sums as above, the first
 * product rounded alone, each later one fused, acc = fma(coef, s, acc).
 * Returns -1 (nothing written) for FP32 stencils with double-formed groups. */
int po_stencil_iterate_fused_f64(const double* in, double* out, int nz, int ny, int nx, const po_stencil* st, int steps);
int po_stencil_iterate_fused_f32(const float* in, float* out, int nz, int ny, int nx, const po_stencil* st, int steps);

/* Reduced-precision storage: fields held as IEEE binary16 ("f16s") or bfloat16
 * ("bf16s") bit patterns, each step computed in float exactly as the f32 path
 * and rounded to the storage format (round to nearest, ties to even). */
uint16_t po_f32_to_f16(float f);
float po_f16_to_f32(uint16_t h);
uint16_t po_f32_to_bf16(float f);
float po_bf16_to_f32(uint16_t b);
void po_stencil_iterate_f16s(const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const po_stencil* st, int steps);
void po_stencil_iterate_bf16s(const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const po_stencil* st, int steps);

/* FP32 CG (float matrix and vectors, dot products in double, same order as
 * po_cg_f64), as the reference's myTest<float,...>.  thr2: the stopping
 * threshold on r.r (already squared; times b.b when relative).  static_iter:
 * run exactly max_iter iterations.  Returns the iterations run. */
int po_cg_f32(int fmt, int n, const int32_t* rowptr, const int32_t* col, const float* val, int ndiag,
              const float* b, float* x, double thr2, int relative, int max_iter, int static_iter,
              double* hist, int hist_cap);

/* Stochastic rounding of 16-bit storage (this library's addition; no
 * reference counterpart).  The random bits of cell `idx` (linear index
 * z*ny*nx + y*nx + x, taken mod 2^32) at time step t (0-based) are
 * lowbias32(idx ^ lowbias32(seed ^ t*0x9E3779B9)), where lowbias32 is
 * C. Wellons' public-domain integer hash.  BF16: (bits(v) + (r & 0xFFFF)) >> 16
 * (NaN -> quiet NaN).  FP16: in the FP16 normal range, the 13 dropped bits get
 * (r & 0x1FFF) added before truncation (results above 65504 -> inf); below it,
 * at zero, inf and NaN, round to nearest even. */
uint32_t po_lowbias32(uint32_t x);
uint32_t po_sr_bits(uint32_t idx, uint32_t step, uint32_t seed);
uint16_t po_f32_to_bf16_sr(float f, uint32_t r);
uint16_t po_f32_to_f16_sr(float f, uint32_t r);
/* fmt: 0 = FP16, 1 = BF16. */
void po_stencil_iterate_sr16(const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const po_stencil* st,
                             int steps, int fmt, uint32_t seed);

/* RMS and max |a-b| over a box [z0,z1)x[y0,y1)x[x0,x1) (common.hpp:133-163). */
double po_rms_error_f64(const double* a, const double* b, int ny, int nx, int z0, int z1, int y0, int y1, int x0, int x1, double* max_err);
double po_rms_error_f32(const float* a, const float* b, int ny, int nx, int z0, int z1, int y0, int y1, int x0, int x1, double* max_err);

/* ---- sparse matrices and CG ----------------------------------------- */
/* 3D Poisson 7-point matrix on an n0 x n1 x n2 grid (x fastest), Dirichlet:
 * diagonal 6, off-diagonals -1, columns ascending.  rowptr has N+1 entries. */
int64_t po_poisson7_nnz(int nx, int ny, int nz);
void po_poisson7_csr(int nx, int ny, int nz, int32_t* rowptr, int32_t* col, double* val);
/* DIA: 7 diagonals with offsets {-nx*ny, -nx, -1, 0, 1, nx, nx*ny}, stored
 * val[d*N + i] = A[i][i+off_d] (0 where i+off_d falls outside or crosses a face). */
void po_poisson7_dia(int nx, int ny, int nz, int32_t* offsets, double* val);

/* CSR SpMV in the library's fixed per-row order.  A row of at most
 * PO_CSR_SHORT_ROW entries is summed in storage order from 0.  A longer row is
 * cut, from its first entry, into chunks of PO_CSR_ROW_CHUNK entries; in a
 * chunk, lane l (0..63) sums entries l, l+64, l+128, ... in order from 0, the
 * 64 lane sums are folded pairwise (v[l] += v[l+s] for s = 32, 16, ..., 1) and
 * v[0] is the chunk's value; the row is chunk 0's value plus each following
 * chunk's value in order.  (The reference sums rows through CUB's merge-path
 * agent, cub/agent/agent_spmv_orig.cuh:666-1100, whose order follows its tile
 * partition; this order depends on the row length only.) */
#define PO_CSR_SHORT_ROW 64
#define PO_CSR_ROW_CHUNK 4096
void po_csr_spmv_f64(int n, const int32_t* rowptr, const int32_t* col, const double* val, const double* x, double* y);
void po_csr_spmv_f32(int n, const int32_t* rowptr, const int32_t* col, const float* val, const float* x, float* y);
void po_dia_spmv_f64(int n, int ndiag, const int32_t* offsets, const double* val, const double* x, double* y);
double po_dot_f64(const double* a, const double* b, size_t n);
double po_dot_f32(const float* a, const float* b, size_t n);

/* FP64 CG following conjugateGradient/cg_kernels.cuh:721-975.  fmt 0 = CSR
 * (rowptr/col/val), 1 = DIA (offsets in `col`, ndiag in `ndiag`).
 * Stops when r.r < tol^2 * (relative ? b.b : 1) or after max_iter iterations.
 * hist[k] = r.r after iteration k (hist[0] = initial), up to hist_cap entries.
 * Returns the iteration count. */
int po_cg_f64(int fmt, int n, const int32_t* rowptr, const int32_t* col, const double* val, int ndiag,
              const double* b, double* x, double tol, int relative, int max_iter,
              double* hist, int hist_cap);

/* FP32-inner / FP64-refinement CG (this build's mixed-precision form; the
 * reference has none).  Outer step: r = b + (-1)*A x in FP64 (x0 = 0 skips
 * the product), stop when r.r < tol^2*(relative ? b.b : 1), after max_outer
 * steps, or when r.r repeats; else an FP32 CG (matrix values and r rounded
 * to float, float SpMV/axpy, dot products of exact double products summed in
 * double as po_dot_f32, alpha/beta = (float)(double ratio)) solves A d = r
 * from d = 0 to r_f.r_f < tau^2 * (r_f0.r_f0), tau^2 = max(inner_tol^2,
 * thr / (4 r.r)) (no further than the outer target needs), and x = x + (double)d.
 * outer_hist[k] = FP64 r.r at outer step k.  Returns total inner iterations;
 * *outer_out = refinement steps taken. */
int po_cg_f32_refine(int fmt, int n, const int32_t* rowptr, const int32_t* col, const double* val, int ndiag,
                     const double* b, double* x, double tol, int relative, double inner_tol, int max_inner,
                     int max_outer, double* outer_hist, int hist_cap, int* outer_out);

#ifdef __cplusplus
}
#endif
#endif
