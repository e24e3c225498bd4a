// jacobi_cuda.hpp -- the reference's C++ entry points, reproduced over the C ABI.
//
// A driver written against the reference's headers
//   stencil/3dstencil/common/jacobi_cuda.cuh:7-12   (j3d_iterative, getMinWidthY)
//   stencil/2dstencil/common/jacobi_cuda.cuh:6-17   (jacobi_iterative)
// includes this header instead and links libperks_amd.so; the calls keep
// their signatures, argument meaning and return codes (0, or the planner's
// -1/-2/-3/-4/-5, or the minimum height when getminHeight is set).
//
// The reference fixes the stencil shape and the kernel variant per
// executable at compile time (config.cuh, -DGEN/-DBASELINE); here the same
// choice is made with two macros, defaulting to the 3d7pt shape with the
// reference's coefficients and clamp boundary, run by the persistent kernel:
//   PERKS_J3D_SHAPE   (enum perks_shape)   PERKS_J3D_METHOD   (enum perks_method)
//   PERKS_J2D_SHAPE   (enum perks_shape)   PERKS_J2D_METHOD   (enum perks_method)
//
// A reference driver builds unmodified against this header when the
// reference's own declaration header is kept out by its include guard:
//   g++ -DPERKS_CUDA_HEADER -include perks_amd/jacobi_cuda.hpp j3d.driver.cpp ...
//
// Reference arguments that select a CUDA kernel variant are accepted and
// recorded in perks_stencil_opts, and do not change what runs here:
//   bdimx         (threads per block, 128/256): the tile shapes are fixed per
//                 kernel variant for gfx950's 64-lane waves (opts.tune[4]);
//   useSM         (LDS plane cache on/off): the kernels keep their working
//                 planes or rows in VGPR rings, LDS only for the row exchange;
//   isDoubleTile  (16 instead of 8 rows per thread): rows per wave are fixed
//                 per tile variant;
//   async         (2D cp.async staging): loads are always asynchronous.
// blkpsm (blocks per SM) caps the workgroups per CU as in the reference;
// usewarmup/warmupiteration run untimed warm-up launches.
#pragma once

#include "../perks_amd.h"

#ifndef PERKS_J3D_SHAPE
#define PERKS_J3D_SHAPE PERKS_SHAPE_3D_STAR_R1
#endif
#ifndef PERKS_J3D_METHOD
#define PERKS_J3D_METHOD PERKS_METHOD_PERSISTENT
#endif
#ifndef PERKS_J2D_SHAPE
#define PERKS_J2D_SHAPE PERKS_SHAPE_2D_STAR_R1
#endif
#ifndef PERKS_J2D_METHOD
#define PERKS_J2D_METHOD PERKS_METHOD_PERSISTENT
#endif

namespace perks_detail {

template <class REAL>
struct Abi;

template <>
struct Abi<double> {
    static constexpr int dtype = PERKS_F64;
    static int j3d(const double* in, int h, int wy, int wx, double* out, int steps, const perks_stencil_opts* o)
    {
        return perks_j3d_iterative_f64(in, h, wy, wx, out, steps, o, nullptr);
    }
    static int j2d(const double* in, int wy, int wx, double* out, int steps, const perks_stencil_opts* o)
    {
        return perks_jacobi2d_iterative_f64(in, wy, wx, out, steps, o, nullptr);
    }
};

template <>
struct Abi<float> {
    static constexpr int dtype = PERKS_F32;
    static int j3d(const float* in, int h, int wy, int wx, float* out, int steps, const perks_stencil_opts* o)
    {
        return perks_j3d_iterative_f32(in, h, wy, wx, out, steps, o, nullptr);
    }
    static int j2d(const float* in, int wy, int wx, float* out, int steps, const perks_stencil_opts* o)
    {
        return perks_jacobi2d_iterative_f32(in, wy, wx, out, steps, o, nullptr);
    }
};

inline perks_stencil_opts make_opts(int shape, int method, int bdimx, int blkpsm, bool useSM, bool usewarmup,
                                    int warmupiteration, bool isDoubleTile)
{
    perks_stencil_opts o;
    perks_stencil_opts_init(&o, shape);
    o.method = method;
    o.bdimx = bdimx;
    o.blkpsm = blkpsm;
    o.use_lds_cache = useSM ? 1 : -1;
    o.double_tile = isDoubleTile ? 1 : 0;
    o.warmup = usewarmup ? 1 : 0;
    o.warmup_steps = warmupiteration > 0 ? warmupiteration : 0;
    return o;
}

}  // namespace perks_detail

// stencil/3dstencil/common/jacobi_cuda.cuh:7 ; j3d.cu:129-140
template <class REAL>
int j3d_iterative(REAL* h_input, int height, int width_y, int width_x, REAL* h_output, int bdimx, int blkpsm,
                  int iteration, bool useSM, bool usewarmup, int warmupiteration, bool isDoubleTile,
                  bool getminHeight = false)
{
    const perks_stencil_opts o = perks_detail::make_opts(PERKS_J3D_SHAPE, PERKS_J3D_METHOD, bdimx, blkpsm, useSM,
                                                         usewarmup, warmupiteration, isDoubleTile);
    if (getminHeight) return perks_j3d_min_height(perks_detail::Abi<REAL>::dtype, width_y, width_x, &o);
    return perks_detail::Abi<REAL>::j3d(h_input, height, width_y, width_x, h_output, iteration, &o);
}

// stencil/3dstencil/common/jacobi_cuda.cuh:12 ; j3d.cu:69-123
template <class REAL>
int getMinWidthY(int width_x, int width_y, int bdimx, bool isDoubleTile = false)
{
    return perks_get_min_width_y(perks_detail::Abi<REAL>::dtype, width_x, width_y, bdimx, isDoubleTile ? 1 : 0);
}

// stencil/2dstencil/common/jacobi_cuda.cuh:14 ; jacobi-2d.cu:137-312
// (registers: the per-thread register budget, 0 = no register cache)
template <class REAL>
int getMinWidthY(int width_x, int bdimx, int registers, bool useSM, int blkpsm, bool isDoubleTile)
{
    return perks_get_min_width_y_2d(perks_detail::Abi<REAL>::dtype, width_x, bdimx, registers, useSM ? 1 : 0, blkpsm,
                                    isDoubleTile ? 1 : 0);
}

// stencil/2dstencil/common/jacobi_cuda.cuh:16 ; jacobi-2d.cu:315-359 (the
// largest over the reference's cache variants: registers and LDS)
template <class REAL>
int getMinWidthY(int width_x, int bdimx, bool isDoubleTile)
{
    return perks_get_min_width_y_2d(perks_detail::Abi<REAL>::dtype, width_x, bdimx, 256, 1, 0, isDoubleTile ? 1 : 0);
}

// stencil/2dstencil/common/jacobi_cuda.cuh:6-10 ; jacobi-2d.cu:369-374
template <class REAL>
int jacobi_iterative(REAL* h_input, int width_y, int width_x, REAL* h_output, int bdimx, int blkpsm, int iteration,
                     bool async, bool useSM, bool usewarmup, int warmupiteration, bool isDoubleTile)
{
    (void)async;  // the reference's cp.async staging flag; loads here are always asynchronous
    const perks_stencil_opts o = perks_detail::make_opts(PERKS_J2D_SHAPE, PERKS_J2D_METHOD, bdimx, blkpsm, useSM,
                                                         usewarmup, warmupiteration, isDoubleTile);
    return perks_detail::Abi<REAL>::j2d(h_input, width_y, width_x, h_output, iteration, &o);
}
