// j3d_common.h -- pieces shared by the 3D Jacobi kernels (plain and persistent).
#pragma once

#include <hip/hip_runtime.h>

#include "perks_amd.h"
#include "perks_device.h"
#include "perks_shapes.h"

namespace perks {

// Coefficients as a kernel argument (in the shape's group order).
struct KCoefs {
    double c[PERKS_MAX_COEFS];
    double divisor;
};

// The 7-point kernels hard-code the neighbour order of kShape3dStarR1
// (E, W, N, S, T, B, C); these checks tie the two together.
constexpr bool star1_order_ok()
{
    const ShapeDesc& s = kShape3dStarR1;
    const int want[7][3] = {{0, 0, 1}, {0, 0, -1}, {0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 0}};
    if (s.ngroups != 7) return false;
    for (int g = 0; g < 7; ++g) {
        if (s.g[g].npts != 1) return false;
        for (int a = 0; a < 3; ++a)
            if (s.g[g].off[0][a] != want[g][a]) return false;
    }
    return true;
}
static_assert(star1_order_ok(), "3D star r1 kernel order differs from the shape table");

// One 7-point update in the reference gold's order (3d7pt_gold_gen.cpp:28-33):
// ((((((cE*E + cW*W) + cN*N) + cS*S) + cT*T) + cB*B) + cC*C), then optional /div.
// Built with -ffp-contract=off, so every * and + rounds separately.
template <typename T, bool HAS_DIV>
__device__ __forceinline__ T star1_update(const T (&k)[8], T e, T w, T n, T s, T t, T b, T c)
{
    T acc = k[0] * e;
    acc = acc + k[1] * w;
    acc = acc + k[2] * n;
    acc = acc + k[3] * s;
    acc = acc + k[4] * t;
    acc = acc + k[5] * b;
    acc = acc + k[6] * c;
    if (HAS_DIV) acc = acc / k[7];
    return acc;
}

// The same sum with each product fused into the running sum (one rounding
// per term instead of two): acc = k0 e, then acc = fma(k_i, x_i, acc) in the
// gold's term order.  Not the gold's bits; within a few ulp of them per step.
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T, bool HAS_DIV>
__device__ __forceinline__ T star1_update_fused(const T (&k)[8], T e, T w, T n, T s, T t, T b, T c)
{
    T acc = k[0] * e;
    acc = fma_t(k[1], w, acc);
    acc = fma_t(k[2], n, acc);
    acc = fma_t(k[3], s, acc);
    acc = fma_t(k[4], t, acc);
    acc = fma_t(k[5], b, acc);
    acc = fma_t(k[6], c, acc);
    if (HAS_DIV) acc = acc / k[7];
    return acc;
}

template <typename T>
__device__ __forceinline__ void load_coefs7(const KCoefs& cf, T (&k)[8])
{
#pragma unroll
    for (int i = 0; i < 7; ++i) k[i] = (T)cf.c[i];
    k[7] = (T)cf.divisor;
}

}  // namespace perks
