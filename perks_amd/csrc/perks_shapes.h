// perks_shapes.h -- stencil shape tables shared by host and device code.
//
// A shape is an ordered list of groups.  Group g contributes
//     coef[g] * (v(p0) + v(p1) + ... )            (points summed left to right)
// and the groups are summed left to right starting from the first group's
// product, followed by an optional division.  The order of groups, the points
// and the reference weights below follow the reference's CPU gold expression
// for each shape, term for term, so that a kernel compiled from this table
// performs the same roundings in the same order:
//
//   3D star r1 (7pt)   stencil/3dstencil/3d7pt/3d7pt_gold_gen.cpp:28-33
//   3D star r2 (13pt)  stencil/3dstencil/3d13pt/3d13pt_gold_gen.cpp (j3d_step expression)
//   3D type0 (17pt)    stencil/3dstencil/3d17pt/3d17pt_gold_gen.cpp:15-35
//   3D box r1 (27pt)   stencil/3dstencil/3d27pt/3d27pt_gold_gen.cpp (j3d_step expression)
//   3D poisson (19pt)  stencil/3dstencil/poisson/poisson_gold_gen.cpp:9-47
//   2D star r1..r6     stencil/2dstencil/{2d5pt,2ds9pt,2d13pt,2d17pt,2d21pt,2ds25pt}/jacobi_gold_gen.cpp
//   2D box r1, r2      stencil/2dstencil/{2d9pt,2d25pt}/jacobi_gold_gen.cpp
//
// Offsets are (dz, dy, dx): +x = east, +y = north, +z = top, matching the
// gold's e0/n0/t0 index names.  `prod_f64` marks a weight the reference writes
// as a double literal (e.g. `15.0*input`): with float storage that product,
// and the running sum from there on, is formed in double.
#pragma once

#include <stdint.h>

#include "perks_amd.h"

namespace perks {

constexpr int kMaxGroups = 27;
constexpr int kMaxPts = 12;

struct ShapeGroup {
    double coef;   // reference weight
    int npts;
    int prod_f64;
    int8_t off[kMaxPts][3];
};

struct ShapeDesc {
    int id;
    int dims;       // 2 or 3
    int radius;     // max |offset| on any axis
    int is_box;     // 1 if any point has two or more non-zero offsets
    int ngroups;
    double divisor; // reference divisor, 0 = none
    ShapeGroup g[kMaxGroups];
};

#define P1(z, y, x) 1, 0, {{z, y, x}}
#define PD(z, y, x) 1, 1, {{z, y, x}}

// clang-format off
constexpr ShapeDesc kShape3dStarR1 = {PERKS_SHAPE_3D_STAR_R1, 3, 1, 0, 7, 0.0, {
    {(double)0.161f, P1(0, 0, 1)},  {(double)0.162f, P1(0, 0, -1)},
    {(double)0.163f, P1(0, 1, 0)},  {(double)0.164f, P1(0, -1, 0)},
    {(double)0.165f, P1(1, 0, 0)},  {(double)0.166f, P1(-1, 0, 0)},
    {(double)-1.67f, P1(0, 0, 0)}}};

constexpr ShapeDesc kShape3dStarR2 = {PERKS_SHAPE_3D_STAR_R2, 3, 2, 0, 13, 0.0, {
    {(double)0.083f, P1(0, 0, 1)},  {(double)0.083f, P1(0, 0, 2)},
    {(double)0.083f, P1(0, 0, -1)}, {(double)0.083f, P1(0, 0, -2)},
    {(double)0.083f, P1(0, 1, 0)},  {(double)0.083f, P1(0, 2, 0)},
    {(double)0.083f, P1(0, -1, 0)}, {(double)0.083f, P1(0, -2, 0)},
    {(double)0.083f, P1(1, 0, 0)},  {(double)0.083f, P1(2, 0, 0)},
    {(double)0.083f, P1(-1, 0, 0)}, {(double)0.083f, P1(-2, 0, 0)},
    {(double)-0.996f, P1(0, 0, 0)}}};

constexpr ShapeDesc kShape3dType0R1 = {PERKS_SHAPE_3D_TYPE0_R1, 3, 1, 1, 11, 159.0, {
    {0.50, 4, 1, {{-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}}},
    {0.51, PD(0, -1, -1)}, {0.71, PD(0, -1, 0)}, {0.91, PD(0, -1, 1)},
    {1.21, PD(0, 0, -1)},  {1.51, PD(0, 0, 0)},  {1.21, PD(0, 0, 1)},
    {0.91, PD(0, 1, -1)},  {0.71, PD(0, 1, 0)},  {0.51, PD(0, 1, 1)},
    {0.52, 4, 1, {{1, -1, -1}, {1, -1, 1}, {1, 1, -1}, {1, 1, 1}}}}};

constexpr ShapeDesc kShape3dBoxR1 = {PERKS_SHAPE_3D_BOX_R1, 3, 1, 1, 27, 159.0, {
    {0.5, PD(-1, -1, -1)},  {0.7, PD(-1, -1, 0)},  {0.9, PD(-1, -1, 1)},
    {1.2, PD(-1, 0, -1)},   {1.5, PD(-1, 0, 0)},   {1.2, PD(-1, 0, 1)},
    {0.9, PD(-1, 1, -1)},   {0.7, PD(-1, 1, 0)},   {0.5, PD(-1, 1, 1)},
    {0.51, PD(0, -1, -1)},  {0.71, PD(0, -1, 0)},  {0.91, PD(0, -1, 1)},
    {1.21, PD(0, 0, -1)},   {1.51, PD(0, 0, 0)},   {1.21, PD(0, 0, 1)},
    {0.91, PD(0, 1, -1)},   {0.71, PD(0, 1, 0)},   {0.51, PD(0, 1, 1)},
    {0.52, PD(1, -1, -1)},  {0.72, PD(1, -1, 0)},  {0.92, PD(1, -1, 1)},
    {1.22, PD(1, 0, -1)},   {1.52, PD(1, 0, 0)},   {1.22, PD(1, 0, 1)},
    {0.92, PD(1, 1, -1)},   {0.72, PD(1, 1, 0)},   {0.52, PD(1, 1, 1)}}};

// 2.666f*c - 0.166f*(6 face neighbours) - 0.0833f*(12 edge neighbours)
constexpr ShapeDesc kShape3dPoissonR1 = {PERKS_SHAPE_3D_POISSON_R1, 3, 1, 1, 3, 0.0, {
    {(double)2.666f, P1(0, 0, 0)},
    {(double)-0.166f, 6, 0, {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, 1}, {0, 0, -1}}},
    {(double)-0.0833f, 12, 0, {{-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {1, 1, 0},
                               {-1, 0, -1}, {-1, 0, 1}, {1, 0, -1}, {1, 0, 1},
                               {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1}}}}};

// 2D stars: far south .. near south, far west .. near west, centre,
// near east .. far east, near north .. far north (the gold's term order).
constexpr ShapeDesc kShape2dStarR1 = {PERKS_SHAPE_2D_STAR_R1, 2, 1, 0, 5, 118.0, {
    {5, P1(0, -1, 0)}, {12, P1(0, 0, -1)}, {15, P1(0, 0, 0)}, {12, P1(0, 0, 1)}, {5, P1(0, 1, 0)}}};

constexpr ShapeDesc kShape2dStarR2 = {PERKS_SHAPE_2D_STAR_R2, 2, 2, 0, 9, 118.0, {
    {7, P1(0, -2, 0)}, {5, P1(0, -1, 0)}, {9, P1(0, 0, -2)}, {12, P1(0, 0, -1)}, {15, P1(0, 0, 0)},
    {12, P1(0, 0, 1)}, {9, P1(0, 0, 2)}, {5, P1(0, 1, 0)}, {7, P1(0, 2, 0)}}};

constexpr ShapeDesc kShape2dStarR3 = {PERKS_SHAPE_2D_STAR_R3, 2, 3, 0, 13, 118.0, {
    {5, P1(0, -3, 0)}, {7, P1(0, -2, 0)}, {5, P1(0, -1, 0)},
    {3, P1(0, 0, -3)}, {9, P1(0, 0, -2)}, {12, P1(0, 0, -1)}, {15, P1(0, 0, 0)},
    {12, P1(0, 0, 1)}, {9, P1(0, 0, 2)}, {3, P1(0, 0, 3)},
    {5, P1(0, 1, 0)}, {7, P1(0, 2, 0)}, {5, P1(0, 3, 0)}}};

constexpr ShapeDesc kShape2dStarR4 = {PERKS_SHAPE_2D_STAR_R4, 2, 4, 0, 17, 118.0, {
    {1, P1(0, -4, 0)}, {5, P1(0, -3, 0)}, {7, P1(0, -2, 0)}, {5, P1(0, -1, 0)},
    {2, P1(0, 0, -4)}, {3, P1(0, 0, -3)}, {9, P1(0, 0, -2)}, {12, P1(0, 0, -1)}, {15, P1(0, 0, 0)},
    {12, P1(0, 0, 1)}, {9, P1(0, 0, 2)}, {3, P1(0, 0, 3)}, {3, P1(0, 0, 4)},
    {5, P1(0, 1, 0)}, {7, P1(0, 2, 0)}, {5, P1(0, 3, 0)}, {4, P1(0, 4, 0)}}};

constexpr ShapeDesc kShape2dStarR5 = {PERKS_SHAPE_2D_STAR_R5, 2, 5, 0, 21, 118.0, {
    {6, P1(0, -5, 0)}, {1, P1(0, -4, 0)}, {5, P1(0, -3, 0)}, {7, P1(0, -2, 0)}, {5, P1(0, -1, 0)},
    {5, P1(0, 0, -5)}, {2, P1(0, 0, -4)}, {3, P1(0, 0, -3)}, {9, P1(0, 0, -2)}, {12, P1(0, 0, -1)},
    {15, P1(0, 0, 0)},
    {12, P1(0, 0, 1)}, {9, P1(0, 0, 2)}, {3, P1(0, 0, 3)}, {3, P1(0, 0, 4)}, {4, P1(0, 0, 5)},
    {5, P1(0, 1, 0)}, {7, P1(0, 2, 0)}, {5, P1(0, 3, 0)}, {4, P1(0, 4, 0)}, {3, P1(0, 5, 0)}}};

constexpr ShapeDesc kShape2dStarR6 = {PERKS_SHAPE_2D_STAR_R6, 2, 6, 0, 25, 118.0, {
    {2, P1(0, -6, 0)}, {6, P1(0, -5, 0)}, {1, P1(0, -4, 0)}, {5, P1(0, -3, 0)}, {7, P1(0, -2, 0)},
    {5, P1(0, -1, 0)},
    {6, P1(0, 0, -6)}, {5, P1(0, 0, -5)}, {2, P1(0, 0, -4)}, {3, P1(0, 0, -3)}, {9, P1(0, 0, -2)},
    {12, P1(0, 0, -1)}, {15, P1(0, 0, 0)}, {12, P1(0, 0, 1)}, {9, P1(0, 0, 2)}, {3, P1(0, 0, 3)},
    {3, P1(0, 0, 4)}, {4, P1(0, 0, 5)}, {6, P1(0, 0, 6)},
    {5, P1(0, 1, 0)}, {7, P1(0, 2, 0)}, {5, P1(0, 3, 0)}, {4, P1(0, 4, 0)}, {3, P1(0, 5, 0)},
    {2, P1(0, 6, 0)}}};

// 2D boxes: rows south to north, columns west to east; the centre weight is
// the double literal 15.0 in the reference.
constexpr ShapeDesc kShape2dBoxR1 = {PERKS_SHAPE_2D_BOX_R1, 2, 1, 1, 9, 118.0, {
    {7, P1(0, -1, -1)}, {5, P1(0, -1, 0)}, {9, P1(0, -1, 1)},
    {12, P1(0, 0, -1)}, {15.0, PD(0, 0, 0)}, {12, P1(0, 0, 1)},
    {9, P1(0, 1, -1)}, {5, P1(0, 1, 0)}, {7, P1(0, 1, 1)}}};

constexpr ShapeDesc kShape2dBoxR2 = {PERKS_SHAPE_2D_BOX_R2, 2, 2, 1, 25, 118.0, {
    {1, P1(0, -2, -2)}, {2, P1(0, -2, -1)}, {3, P1(0, -2, 0)}, {4, P1(0, -2, 1)}, {5, P1(0, -2, 2)},
    {7, P1(0, -1, -2)}, {7, P1(0, -1, -1)}, {5, P1(0, -1, 0)}, {7, P1(0, -1, 1)}, {6, P1(0, -1, 2)},
    {8, P1(0, 0, -2)}, {12, P1(0, 0, -1)}, {15.0, PD(0, 0, 0)}, {12, P1(0, 0, 1)}, {12, P1(0, 0, 2)},
    {9, P1(0, 1, -2)}, {9, P1(0, 1, -1)}, {5, P1(0, 1, 0)}, {7, P1(0, 1, 1)}, {15, P1(0, 1, 2)},
    {10, P1(0, 2, -2)}, {11, P1(0, 2, -1)}, {12, P1(0, 2, 0)}, {13, P1(0, 2, 1)}, {14, P1(0, 2, 2)}}};
// clang-format on

#undef P1
#undef PD

// Host-side lookup (nullptr for an unknown id).
inline const ShapeDesc* shape_by_id(int id)
{
    switch (id) {
    case PERKS_SHAPE_3D_STAR_R1: return &kShape3dStarR1;
    case PERKS_SHAPE_3D_STAR_R2: return &kShape3dStarR2;
    case PERKS_SHAPE_3D_TYPE0_R1: return &kShape3dType0R1;
    case PERKS_SHAPE_3D_BOX_R1: return &kShape3dBoxR1;
    case PERKS_SHAPE_3D_POISSON_R1: return &kShape3dPoissonR1;
    case PERKS_SHAPE_2D_STAR_R1: return &kShape2dStarR1;
    case PERKS_SHAPE_2D_STAR_R2: return &kShape2dStarR2;
    case PERKS_SHAPE_2D_STAR_R3: return &kShape2dStarR3;
    case PERKS_SHAPE_2D_STAR_R4: return &kShape2dStarR4;
    case PERKS_SHAPE_2D_STAR_R5: return &kShape2dStarR5;
    case PERKS_SHAPE_2D_STAR_R6: return &kShape2dStarR6;
    case PERKS_SHAPE_2D_BOX_R1: return &kShape2dBoxR1;
    case PERKS_SHAPE_2D_BOX_R2: return &kShape2dBoxR2;
    default: return nullptr;
    }
}

}  // namespace perks
