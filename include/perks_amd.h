/*
 * perks_amd.h -- C ABI of the MI355X (gfx950) persistent Jacobi stencil / CG library.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no HIP or torch
 * types.  `stream` arguments take a hipStream_t cast to void* (NULL = the
 * default stream).  Every function returns PERKS_OK (0) or a negative error
 * code; nothing in the library calls exit().
 *
 * Reference interfaces these entry points replace (paths in the PERKS tree):
 *   perks_j3d_iterative_{f64,f32}      template<class REAL> int j3d_iterative(...)
 *                                      stencil/3dstencil/common/jacobi_cuda.cuh:7-8,
 *                                      defined stencil/3dstencil/j3d.cu:129-808
 *   perks_j3d_min_height               j3d_iterative(..., getminHeight=true)
 *                                      stencil/3dstencil/j3d.cu:598,612
 *   perks_get_min_width_y              template<class REAL> int getMinWidthY(int,int,int,bool)
 *                                      stencil/3dstencil/common/jacobi_cuda.cuh:12; j3d.cu:69-123
 *   perks_get_min_width_y_2d           template<class REAL> int getMinWidthY(int,int,int,bool,int,bool)
 *                                      and getMinWidthY(int,int,bool)
 *                                      stencil/2dstencil/common/jacobi_cuda.cuh:12-17; jacobi-2d.cu:79-364
 *   perks_jacobi2d_iterative_{f64,f32} template<class REAL> int jacobi_iterative(...)
 *                                      stencil/2dstencil/common/jacobi_cuda.cuh:6-10;
 *                                      stencil/2dstencil/jacobi-2d.cu:369-865
 *   perks_cg_solve_f64                 myTest<double,int,...> + DispatchCG::ProcessDispatchCG
 *                                      conjugateGradient/cg_driver.cuh:4-5; cg_driver.cu:53-428;
 *                                      conjugateGradient/cg_dispatcher.cuh:559-916
 * The reference's C++ template signatures are reproduced on top of this ABI
 * in include/perks_amd/jacobi_cuda.hpp.
 */
#ifndef PERKS_AMD_H
#define PERKS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (negative values keep the reference planner's meaning) -- */
#define PERKS_OK 0
#define PERKS_ERR_NO_LDS_FOR_CACHE (-1) /* j3d.cu:591  useSM but no LDS plane fits      */
#define PERKS_ERR_UNSUPPORTED_SIZE (-2) /* j3d.driver.cpp:186  x/y widths not supported */
#define PERKS_ERR_NOT_LAUNCHABLE (-3)   /* j3d.cu:570  occupancy is 0                   */
#define PERKS_ERR_HEIGHT_TOO_SMALL (-4) /* j3d.cu:606  gridDim.z*(2H+1) > height         */
#define PERKS_ERR_SPILL (-5)            /* j3d.cu:482  cache depth would spill          */
#define PERKS_ERR_INVALID_ARG (-10)
#define PERKS_ERR_HIP (-11)            /* a HIP runtime call failed                    */
#define PERKS_ERR_NO_DEVICE (-12)
#define PERKS_ERR_UNSUPPORTED (-13)     /* shape/dtype/method combination not built     */
#define PERKS_ERR_NOT_CONVERGED (-14)
#define PERKS_ERR_BARRIER_TIMEOUT (-15) /* a grid barrier wait exceeded 10 s; launch abandoned */

/* ---- enums ---------------------------------------------------------------- */
enum perks_rounding { PERKS_ROUND_NEAREST = 0, PERKS_ROUND_STOCHASTIC = 1 };
enum perks_dtype { PERKS_F64 = 0, PERKS_F32 = 1, PERKS_F16 = 2, PERKS_BF16 = 3 };

/* Shapes are named by (dimension, kind, radius).  Reference directory names:
 * 3d7pt, 3d13pt, 3d17pt, 3d27pt, poisson; 2d5pt, 2ds9pt, 2d13pt, 2d17pt,
 * 2d21pt, 2ds25pt (stars r1..r6); 2d9pt, 2d25pt (boxes r1, r2). */
enum perks_shape {
    PERKS_SHAPE_3D_STAR_R1 = 0,
    PERKS_SHAPE_3D_STAR_R2 = 1,
    PERKS_SHAPE_3D_TYPE0_R1 = 2,
    PERKS_SHAPE_3D_BOX_R1 = 3,
    PERKS_SHAPE_3D_POISSON_R1 = 4,
    PERKS_SHAPE_2D_STAR_R1 = 10,
    PERKS_SHAPE_2D_STAR_R2 = 11,
    PERKS_SHAPE_2D_STAR_R3 = 12,
    PERKS_SHAPE_2D_STAR_R4 = 13,
    PERKS_SHAPE_2D_STAR_R5 = 14,
    PERKS_SHAPE_2D_STAR_R6 = 15,
    PERKS_SHAPE_2D_BOX_R1 = 16,
    PERKS_SHAPE_2D_BOX_R2 = 17
};

/* Boundary: CLAMP replicates the edge cell and updates every cell (the
 * reference's rule, 3d7pt_gold_gen.cpp:21-26).  ZERO treats the array as the
 * interior of a grid whose one-cell (radius-cell) halo is fixed at 0. */
enum perks_boundary { PERKS_BC_CLAMP = 0, PERKS_BC_ZERO = 1 };

/* PLAIN: one kernel launch per time step (the reference's kernel3d_baseline).
 * PERSISTENT: one cooperative launch runs every step (kernel3d_general). */
enum perks_method { PERKS_METHOD_PLAIN = 0, PERKS_METHOD_PERSISTENT = 1 };

#define PERKS_MAX_COEFS 27

typedef struct perks_stencil_opts {
    int32_t shape;          /* enum perks_shape */
    int32_t boundary;       /* enum perks_boundary */
    int32_t method;         /* enum perks_method */
    int32_t use_ref_coefs;  /* 1: the shape's reference weights and divisor */
    double coef[PERKS_MAX_COEFS]; /* per group, in the shape's summation order */
    double divisor;         /* 0 = none (ignored when use_ref_coefs) */
    /* tuning knobs; 0 = automatic.  Names follow j3d.driver.cpp:45-57. */
    int32_t bdimx;          /* --bdim: 128 or 256 threads per workgroup */
    int32_t blkpsm;         /* --blkpsm: cap on workgroups per CU */
    int32_t use_lds_cache;  /* --usesm: 1 on, 0 auto, -1 off */
    int32_t reg_cache_planes; /* VGPR-cached planes per workgroup; -1 auto */
    int32_t double_tile;    /* --doubletile */
    int32_t warmup;         /* --warmup: run untimed launches first */
    int32_t warmup_steps;   /* --warmiter */
    /* experimental knobs, 0 = automatic:
     *   tune[0] plain: z-chunk per workgroup       tune[1] plain: 2 = cached stores (default nontemporal)
     *   tune[2] persistent: time steps per pass     tune[3] persistent: workgroups per CU
     *           (3D 7-point default: 4; under the clamp boundary 3 for FP16 and for BF16
     *           with nearest rounding; perks_j3d_autotune measures every choice)
     *   tune[4] tile variant.  3D 7-point persistent: 0..3 (rows x waves of its kernel;
     *           FP64 0 = 12 waves x 3 rows, 3 = 8 waves x 5 rows), 4 = two columns per lane (FP32/FP16/BF16, even
     *           widths) or, for FP64, 8 waves x 4 rows; 5 = resident tiles (FP64/FP32): the whole domain held
     *           in VGPRs across every step, one grid barrier per step, halos only through
     *           memory; PERKS_ERR_UNSUPPORTED_SIZE above perks_j3d_min_height's height.
     *           3D other shapes: 0 = default, 2 = register streaming (64 x 8 tiles),
     *           4 = register streaming (64 x 16 tiles), 3 = LDS planes.
     *           2D: 0 = default (4 where it applies, else 2), 1 = LDS tile 128x64 (64x64 FP64),
     *           2 = row streaming through registers, one column per lane, 3 = LDS tile 128x32
     *           (64x32 FP64), 4 = row streaming, two columns per lane (FP32/FP16/BF16 storage,
     *           even widths; otherwise PERKS_ERR_UNSUPPORTED)
     *   tune[5] 3D: 1 = run the 7-point star on the shape-generic kernel
     *   tune[6] 3D 7-point persistent, FP64/FP32 storage, no divisor: 1 = contracted
     *           update (the gold's term order with each product fused into the running
     *           sum by an FMA; not the gold's bits, within a few ulp of them per step);
     *           PERKS_ERR_UNSUPPORTED for any other method, shape, format or a divisor.
     *           0 (default) = uncontracted, bit-identical to the reference CPU gold
     *   tune[7] 3D 7-point persistent (ring kernel): z-slabs per x-y tile, each one a
     *           work item of a pass (capped at the height); 0 = the planner's choice */
    int32_t tune[8];
    /* FP16/BF16 storage only (no reference counterpart): how each step's FP32
     * result is rounded to the storage format.  PERKS_ROUND_NEAREST (default):
     * nearest even.  PERKS_ROUND_STOCHASTIC: up or down with probability
     * proportional to the distance, the random bits being a counter hash of
     * (cell index, time step, round_seed), so runs are reproducible and the
     * per-step and persistent methods agree bit for bit.  Nearest rounding
     * drops every change smaller than half a unit in the last place, so a
     * slowly decaying field stalls; stochastic rounding keeps it unbiased. */
    int32_t rounding;
    uint32_t round_seed;
} perks_stencil_opts;

typedef struct perks_run_report {
    double elapsed_ms;     /* device time of the launch(es), hipEvents on `stream` */
    double gcells_per_s;   /* steps * cells / elapsed (j3d.cu:754)               */
    int32_t grid[3];
    int32_t block;
    int32_t lds_bytes;
    int32_t lds_planes;    /* planes kept in LDS per workgroup (persistent)      */
    int32_t reg_planes;    /* planes kept in VGPRs per workgroup (persistent)    */
    int32_t result_in_b;   /* device entry points: 1 if the result is in d_b     */
    int32_t occupancy;     /* workgroups per CU used                            */
    int32_t reserved[7];
} perks_run_report;

/* Device facts the planner needs; perks_query_device fills it from HIP. */
typedef struct perks_device_caps {
    int32_t cu_count;
    int32_t lds_per_cu;      /* bytes */
    int32_t lds_per_block;   /* bytes */
    int32_t vgprs_per_simd;  /* 32-bit registers per lane per SIMD */
    int32_t simds_per_cu;
    int32_t warp_size;
    int32_t cooperative;
    int32_t reserved[5];
} perks_device_caps;

typedef struct perks_plan_info {
    int32_t grid[3];
    int32_t block;
    int32_t occupancy;       /* workgroups per CU                   */
    int32_t lds_bytes;
    int32_t lds_planes;
    int32_t reg_planes;
    int32_t min_height;      /* getminHeight result (j3d.cu:598)    */
    int32_t tile_x, tile_y;
    int32_t reserved[5];
} perks_plan_info;

/* ---- library ------------------------------------------------------------- */
const char* perks_version(void);
const char* perks_error_string(int code);
int perks_query_device(int device, perks_device_caps* caps);
/* Frees cached device scratch of the current device. */
int perks_release(void);

/* ---- shapes --------------------------------------------------------------- */
/* Fills opts with defaults for `shape`: reference weights, CLAMP, PERSISTENT. */
void perks_stencil_opts_init(perks_stencil_opts* opts, int shape);
/* Describes a shape.  Any output pointer may be NULL.  coef/npts/prod_f64
 * hold *ngroups entries; offsets holds ngroups*12*3 int8 (dz,dy,dx). */
int perks_shape_info(int shape, int32_t* dims, int32_t* radius, int32_t* ngroups, double* coef,
                     double* divisor, int32_t* npts, int32_t* prod_f64, int8_t* offsets);

/* ---- planning (host only when caps != NULL) ------------------------------ */
int perks_j3d_plan(int dtype, int height, int width_y, int width_x, const perks_stencil_opts* opts,
                   const perks_device_caps* caps, int occupancy_hint, perks_plan_info* out);
/* j3d_iterative(..., getminHeight=true): the largest height whose domain the
 * persistent kernel holds on chip for the whole solve (7-point star, FP64/FP32:
 * the resident-tile variant, opts.tune[4] = 5; otherwise the height whose
 * field fits the chip's LDS + VGPRs). */
int perks_j3d_min_height(int dtype, int width_y, int width_x, const perks_stencil_opts* opts);
/* getMinWidthY<REAL>(width_x, width_y, bdimx, isDoubleTile), j3d.cu:69-123. */
int perks_get_min_width_y(int dtype, int width_x, int width_y, int bdimx, int double_tile);
/* The 2D getMinWidthY overloads (stencil/2dstencil/common/jacobi_cuda.cuh:12-17;
 * defined stencil/2dstencil/jacobi-2d.cu:79-364): the number of rows of
 * width_x cells that the chip's on-chip storage holds, i.e. the height below
 * which a persistent grid could keep the whole 2D domain resident.
 * `registers` > 0 counts every CU's VGPR file, use_lds every CU's LDS; as in
 * the reference, no cache at all (registers == 0 and !use_lds) returns 0.
 * blkpsm and double_tile do not change the capacity and are accepted for the
 * signature.  The reference's 3-argument form is registers = 256, use_lds = 1
 * (its maximum over the cache variants). */
int perks_get_min_width_y_2d(int dtype, int width_x, int bdimx, int registers, int use_lds, int blkpsm,
                             int double_tile);

/* ---- 3D Jacobi ------------------------------------------------------------ */
/* Host buffers, as the reference: h_in[height][width_y][width_x] (x fastest)
 * is read, h_out receives the field after `steps` steps.  The callee
 * allocates and frees all device memory.  report may be NULL. */
int perks_j3d_iterative_f64(const double* h_in, int height, int width_y, int width_x, double* h_out,
                            int steps, const perks_stencil_opts* opts, perks_run_report* report);
int perks_j3d_iterative_f32(const float* h_in, int height, int width_y, int width_x, float* h_out,
                            int steps, const perks_stencil_opts* opts, perks_run_report* report);
/* Device-resident buffers: d_a holds the input field, d_b is a second field
 * of the same size.  Both are overwritten; after return (the work is
 * enqueued on `stream` and the call waits for it), the result is in d_b
 * when report->result_in_b is 1 (odd step counts), else in d_a. */
int perks_j3d_device_f64(double* d_a, double* d_b, int height, int width_y, int width_x, int steps,
                         const perks_stencil_opts* opts, void* stream, perks_run_report* report);
int perks_j3d_device_f32(float* d_a, float* d_b, int height, int width_y, int width_x, int steps,
                         const perks_stencil_opts* opts, void* stream, perks_run_report* report);
/* Any storage type (enum perks_dtype): FP64, FP32, or FP16/BF16 storage (raw
 * 16-bit patterns) with FP32 arithmetic, every step's result rounded to the
 * storage format (nearest even).  The per-step method needs an even width_x
 * for 16-bit storage (PERKS_ERR_UNSUPPORTED_SIZE otherwise).  No reference
 * counterpart: the reduced-precision form is an addition of this library. */
int perks_j3d_device(int dtype, void* d_a, void* d_b, int height, int width_y, int width_x, int steps,
                     const perks_stencil_opts* opts, void* stream, perks_run_report* report);

/* ---- 3D configuration table and autotuner ---------------------------------- */
/* One launch configuration of the 3D path for a problem (method, tile
 * variant, time steps per pass) and what its loaded code object reports.
 * Replaces one row of the reference's register-folder table
 * (stencil/3dstencil/perksconfig.cuh, regfolder<>), which the reference
 * derives at build time from nvcc's register and spill report
 * (ProcessCompilerLog.py:47-75, ProcessRst2cuh.py:5-51). */
typedef struct perks_kernel_config {
    int32_t method;        /* PERKS_METHOD_PLAIN or PERKS_METHOD_PERSISTENT       */
    int32_t variant;       /* opts.tune[4]                                        */
    int32_t levels;        /* opts.tune[2]: time steps per pass (1 for plain)     */
    int32_t block;         /* threads per workgroup                               */
    int32_t vgprs;         /* architected VGPRs per lane                          */
    int32_t spill_bytes;   /* scratch bytes per lane; nonzero means spills        */
    int32_t lds_bytes;     /* LDS per workgroup                                   */
    int32_t occupancy;     /* workgroups per CU the launch would use              */
    int32_t tile_x;        /* x-y tile of one work item                           */
    int32_t tile_y;
    int32_t status;        /* PERKS_OK, or why the configuration does not apply   */
    int32_t reserved0;
    double gcells_per_s;   /* measured by perks_j3d_autotune (0 from the table)   */
    int32_t reserved[4];
} perks_kernel_config;

/* Lists every configuration the library holds for this problem (dtype,
 * sizes, opts' shape, boundary, coefficients and rounding), writing at most
 * `capacity` entries and the total to *count.  Needs the device (attribute
 * and occupancy queries) but launches nothing. */
int perks_j3d_kernel_table(int dtype, int height, int width_y, int width_x, const perks_stencil_opts* opts,
                           perks_kernel_config* table, int capacity, int* count);
/* Times every configuration of the table whose status is PERKS_OK on
 * `sample_steps` steps (<= 0: 12, a multiple of every pass depth) of the
 * field d_in (left unchanged; the runs use two scratch fields the call
 * allocates and frees), then writes the fastest into opts: method, tune[2]
 * and tune[4].  Every configuration computes the same bits, so the choice
 * changes only the speed.  table/count as above (table may be NULL). */
int perks_j3d_autotune(int dtype, const void* d_in, int height, int width_y, int width_x, int sample_steps,
                       perks_stencil_opts* opts, perks_kernel_config* table, int capacity, int* count,
                       void* stream);

/* ---- 2D Jacobi ------------------------------------------------------------ */
/* Host buffers h_in[width_y][width_x] (x fastest), as the reference's
 * jacobi_iterative<REAL>.  report may be NULL. */
int perks_jacobi2d_iterative_f64(const double* h_in, int width_y, int width_x, double* h_out, int steps,
                                 const perks_stencil_opts* opts, perks_run_report* report);
int perks_jacobi2d_iterative_f32(const float* h_in, int width_y, int width_x, float* h_out, int steps,
                                 const perks_stencil_opts* opts, perks_run_report* report);
/* Device-resident field of any storage type (PERKS_F64, PERKS_F32, and
 * PERKS_F16 / PERKS_BF16 storage with FP32 arithmetic).  d_a holds the
 * input; d_a and d_b are both overwritten; the result is in d_b when
 * report->result_in_b is 1, else in d_a. */
int perks_jacobi2d_device(int dtype, void* d_a, void* d_b, int width_y, int width_x, int steps,
                          const perks_stencil_opts* opts, void* stream, perks_run_report* report);

/* Batched 2D Jacobi: many independent grids of any sizes and step counts in
 * one persistent launch (PERKS_METHOD_PERSISTENT) or as per-step launches
 * (PERKS_METHOD_PLAIN).  All grids share dtype, shape and coefficients.  On
 * return each grid's result_in_b tells where its result is. */
typedef struct perks_grid2d {
    void* a;
    void* b;
    int32_t width_y, width_x;
    int32_t steps;
    int32_t result_in_b; /* out */
} perks_grid2d;
int perks_jacobi2d_batch(int dtype, perks_grid2d* grids, int ngrids, const perks_stencil_opts* opts, void* stream,
                         perks_run_report* report);

/* perks_jacobi2d_batch without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  The launch's scratch --
 * grid-barrier block, work list, per-pass active counts -- lives in a fixed
 * device array (about 0.9 MiB) that each GPU holds until the process ends,
 * so a batch may hold at most PERKS_NOALLOC_MAX_GRIDS grids with steps and
 * PERKS_NOALLOC_MAX_PASSES passes (per-step method: time steps) of its
 * longest grid; larger batches return PERKS_ERR_UNSUPPORTED_SIZE.  Calls on
 * one GPU share that array: calls from one process are serialised, and the
 * call waits for its work before returning.  report->elapsed_ms is 0 (no
 * timing events are created).  Results equal perks_jacobi2d_batch's. */
#define PERKS_NOALLOC_MAX_GRIDS 4096
#define PERKS_NOALLOC_MAX_PASSES 16384
int perks_jacobi2d_batch_noalloc(int dtype, perks_grid2d* grids, int ngrids, const perks_stencil_opts* opts,
                                 void* stream, perks_run_report* report);

/* A reusable plan for a batch (no reference counterpart: the batched form is
 * an addition of this library).  Creating it uploads the work list once;
 * with autotune = 1 it times candidate launch configurations (tile shape,
 * time steps per pass, workgroups per CU) on scratch buffers of the batch's
 * sizes (at most 4 GiB of them, allocated and freed during the call) and keeps
 * the fastest; then it captures the launch in a hipGraph.  Each
 * perks_jacobi2d_batch_run replays the graph on `stream` and waits: every grid
 * runs its steps from what its buffer a holds, as perks_jacobi2d_batch does.
 * The buffers, sizes and step counts are fixed by the plan.  A plan owns its
 * grid-barrier block (perks_release() leaves it valid); runs of one plan are
 * serialised.  The captured launch is an ordinary (not cooperative) launch of
 * at most CUs x occupancy workgroups whose barrier needs every workgroup
 * resident at once: a replay needs the device to itself, and other work that
 * holds CUs meanwhile makes its barrier wait time out (PERKS_ERR_BARRIER_TIMEOUT
 * after 10 s) instead of completing.
 * plan_info: result_in_b[i] per grid (ngrids entries, may be NULL) and
 * config[6] = {tile variant, steps per pass, workgroups per CU, workgroups,
 * autotuned, passes} (may be NULL). */
typedef struct perks_batch_plan perks_batch_plan;
int perks_jacobi2d_batch_plan(int dtype, const perks_grid2d* grids, int ngrids, const perks_stencil_opts* opts,
                              int autotune, void* stream, perks_batch_plan** plan);
int perks_jacobi2d_batch_run(perks_batch_plan* plan, void* stream, perks_run_report* report);
int perks_jacobi2d_batch_plan_info(const perks_batch_plan* plan, int32_t* result_in_b, int ngrids, int32_t* config);
void perks_jacobi2d_batch_plan_destroy(perks_batch_plan* plan);

/* ---- conjugate gradient ---------------------------------------------------- */
/* Device-resident CSR: rowptr[n+1], col[nnz], val[nnz] (columns ascending
 * within a row, as the reference's CSR built by sparse_matrix.h:666-734). */
typedef struct perks_csr {
    int32_t n;
    int32_t reserved0;
    int64_t nnz;
    const int32_t* rowptr;
    const int32_t* col;
    const double* val;
} perks_csr;

/* DIA: ndiag (<= 16) diagonals; offsets is a HOST array, val a device array
 * of ndiag*n values with val[d*n + i] = A[i][i + offsets[d]] (entries whose
 * column falls outside [0,n) are ignored). */
typedef struct perks_dia {
    int32_t n;
    int32_t ndiag;
    const int32_t* offsets;
    const double* val;
} perks_dia;

enum perks_cg_precision { PERKS_CG_F64 = 0, PERKS_CG_F32_REFINE = 1 };

typedef struct perks_cg_opts {
    double tol;          /* stop when r.r < tol^2 (* b.b when relative)      */
    double inner_tol;    /* F32_REFINE: relative tolerance of each inner solve; a solve stops
                            earlier, at sqrt(tol^2 / r.r) / 2, when that is enough for tol */
    int32_t relative;    /* 1: ||r||/||b|| < tol ; 0: ||r|| < tol (cg_kernels.cuh:834) */
    int32_t max_iter;    /* iterations (F64) / per inner solve (F32_REFINE) */
    int32_t precision;   /* enum perks_cg_precision */
    int32_t max_inner;   /* F32_REFINE: cap per inner solve (0 = max_iter) */
    int32_t max_outer;   /* F32_REFINE: refinement steps (0 = 20)         */
    int32_t x0_is_zero;  /* 1: skip the initial A x (x must hold zeros)    */
    int32_t blkpsm;      /* cap on workgroups per CU (0 = occupancy)       */
    int32_t static_iter; /* 1: run exactly max_iter iterations (--staticiter, cg_kernels.cuh:834) */
    int32_t cache_vector;/* 0: keep the workgroups' own rows of Ap, x, p, r in LDS across
                            iterations while LDS lasts (cacheVector); -1: off            */
    int32_t cache_matrix;/* 0: keep the leading matrix entries (CSR) or rows (DIA) of each
                            workgroup in LDS with what LDS is left (cacheMatrix); -1: off */
    int32_t reserved[2];
} perks_cg_opts;

typedef struct perks_cg_report {
    int32_t iterations;       /* CG iterations (summed over inner solves)    */
    int32_t outer_iterations; /* refinement steps (F32_REFINE)              */
    double final_rr;          /* r.r of the FP64 residual at exit            */
    double bb;                /* b.b                                         */
    double elapsed_ms;        /* device time of the solve (hipEvents)        */
    int32_t grid;             /* workgroups of the persistent launch         */
    int32_t hist_cap;         /* capacity of history (caller-provided)       */
    double* history;          /* host array: r.r after each iteration (F64:   */
                              /* [0] initial) or each refinement step        */
    int32_t hist_len;         /* entries produced (may exceed hist_cap)      */
    int32_t lds_bytes;        /* LDS cache per workgroup (bytes)             */
    int32_t cache_units;      /* 256-row vector units cached per workgroup   */
    int32_t cache_matrix;     /* CSR entries / DIA rows cached per workgroup */
    int32_t long_rows;        /* CSR rows split across the grid (> 64 entries) */
    int32_t reserved[1];
} perks_cg_report;

void perks_cg_opts_init(perks_cg_opts* opts);
/* Solves A x = b; d_b, d_x device vectors of n doubles (x is the initial
 * guess and receives the solution).  Returns PERKS_OK when converged,
 * PERKS_ERR_NOT_CONVERGED when max_iter ran out (x and report are valid). */
int perks_cg_solve_csr(const perks_csr* A, const double* d_b, double* d_x, const perks_cg_opts* opts, void* stream,
                       perks_cg_report* report);
int perks_cg_solve_dia(const perks_dia* A, const double* d_b, double* d_x, const perks_cg_opts* opts, void* stream,
                       perks_cg_report* report);

/* FP32 CG (the reference's myTest<float,...>): FP32 matrix and vectors, dot
 * products accumulated in double as the reference's CgParams<float, double>
 * does; the stopping threshold is (float)tol * (float)tol (times b.b when
 * relative).  opts->precision is ignored. */
typedef struct perks_csr_f32 {
    int32_t n;
    int32_t reserved0;
    int64_t nnz;
    const int32_t* rowptr;
    const int32_t* col;
    const float* val;
} perks_csr_f32;
typedef struct perks_dia_f32 {
    int32_t n;
    int32_t ndiag;
    const int32_t* offsets; /* host array */
    const float* val;
} perks_dia_f32;
int perks_cg_solve_csr_f32(const perks_csr_f32* A, const float* d_b, float* d_x, const perks_cg_opts* opts,
                           void* stream, perks_cg_report* report);
int perks_cg_solve_dia_f32(const perks_dia_f32* A, const float* d_b, float* d_x, const perks_cg_opts* opts,
                           void* stream, perks_cg_report* report);

/* ---- the reference's CG driver -------------------------------------------------
 * myTest<ValueT,int,...> (conjugateGradient/cg_driver.cu:53-428): builds the
 * matrix (a Matrix Market file, or the reference's virtual random tridiagonal
 * matrix of vdata_n rows), sets rhs = 1 and x = 0, runs the persistent CG in
 * FP64 (value_bytes 8) or FP32 (4) to ||r|| < tol (absolute, as the
 * reference), optionally checks max|A x - rhs| on the host, and prints the
 * reference's report lines when verbose.  Device memory is allocated and
 * freed inside the call. */
typedef struct perks_cg_driver_opts {
    const char* mtx_path;  /* NULL or "": the virtual tridiagonal matrix (--vdata) */
    int32_t value_bytes;   /* 8: double, 4: float                              */
    int32_t vdata_n;       /* rows of the virtual matrix (0: 1048576)          */
    int32_t max_iter;      /* --iters; <= 0: 10 * rows (file) or rows (virtual) */
    int32_t static_iter;   /* --staticiter                                     */
    int32_t warmup;        /* --warmup: one untimed 2-iteration solve first    */
    int32_t check;         /* --check: max |A x - rhs| on the host             */
    int32_t verbose;       /* print the reference's report lines              */
    int32_t reserved0;
    double tol;            /* 0: 1e-10                                         */
} perks_cg_driver_opts;
typedef struct perks_cg_driver_result {
    int32_t n;
    int32_t iterations;
    int64_t nnz;
    double residual;       /* sqrt(r.r) at exit                                */
    double elapsed_ms;     /* device time of the timed solve                   */
    double ms_per_iter;
    double gbps;           /* the reference's byte model per iteration / time  */
    double check_error;    /* max |A x - rhs| when check, else -1              */
    int32_t grid;
    int32_t code;          /* PERKS_OK or PERKS_ERR_NOT_CONVERGED              */
} perks_cg_driver_result;
int perks_cg_driver(const perks_cg_driver_opts* opts, perks_cg_driver_result* result);

/* ---- sparse inputs (host) ----------------------------------------------------
 * Matrix Market reader and COO -> CSR as the reference's CooMatrix::InitMarket /
 * CsrMatrix::Init (conjugateGradient/matrix/sparse_matrix.h:217-390, 666-734):
 * entries stable-sorted by (row, column), duplicates kept, symmetric files
 * mirrored (skew: negated), pattern entries 1, array files column-major.
 * The arrays are allocated by the library; free each with perks_free_host. */
int perks_mm_read_csr(const char* path, int32_t* n, int64_t* nnz, int32_t** rowptr, int32_t** col, double** val);
void perks_free_host(void* p);
/* The reference's random symmetric tridiagonal matrix (conjugateGradient/
 * matrix/tridiag.cpp:11-47, glibc rand()): n rows, nnz = 3n - 2, arrays of
 * n+1 / nnz / nnz entries supplied by the caller. */
int perks_gen_tridiag_f64(int32_t n, int32_t* rowptr, int32_t* col, double* val);
int perks_gen_tridiag_f32(int32_t n, int32_t* rowptr, int32_t* col, float* val);
/* CSR -> DIA on the host: offsets (ascending) of the diagonals that hold
 * entries; dia_val (may be NULL) receives ndiag*n values (duplicate entries of
 * one position added).  Returns ndiag, or PERKS_ERR_UNSUPPORTED when more
 * than max_diag diagonals are needed. */
int perks_csr_to_dia(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, int32_t max_diag,
                     int32_t* offsets, double* dia_val);

/* 3D Poisson 7-point operator on nx*ny*nz points (x fastest, Dirichlet):
 * diagonal 6, off-diagonals -1.  CSR arrays and the 7 DIA diagonals (offsets
 * {-nx*ny, -nx, -1, 0, 1, nx, nx*ny}) are generated on the device. */
int64_t perks_poisson7_nnz(int nx, int ny, int nz);
int perks_poisson7_csr(int nx, int ny, int nz, int32_t* d_rowptr, int32_t* d_col, double* d_val, void* stream);
int perks_poisson7_dia(int nx, int ny, int nz, int32_t* offsets, double* d_val, void* stream);

/* ---- inputs ---------------------------------------------------------------- */
/* d[i] = lo + (hi-lo) * u53(splitmix64(seed + (i+1)*0x9E3779B97F4A7C15)),
 * the same sequence the host generator produces (float: rounded once). */
int perks_fill_uniform_f64(double* d, size_t n, uint64_t seed, double lo, double hi, void* stream);
int perks_fill_uniform_f32(float* d, size_t n, uint64_t seed, double lo, double hi, void* stream);
/* The f32 sequence above, each value then rounded to binary16 / bfloat16
 * (nearest even). */
int perks_fill_uniform_f16(uint16_t* d, size_t n, uint64_t seed, double lo, double hi, void* stream);
int perks_fill_uniform_bf16(uint16_t* d, size_t n, uint64_t seed, double lo, double hi, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PERKS_AMD_H */
