// j3d_general_kernel.h -- shape-generic 3D Jacobi kernel (every 3D shape of the
// reference: 7- and 13-point stars, the 17-point type-0 stencil, the 27-point
// box and the 19-point Poisson stencil), for FP64, FP32 and FP16/BF16 storage.
//
// A workgroup owns an x-y tile of TX x TY cells (a ghost rim of R*L cells
// included, R = shape radius, L = time steps in the pass) over a z-slab and
// marches along z.  Each level l = 0..L-1 keeps a ring of 2R+1 planes of the
// tile in LDS; at iteration i the workgroup loads level-0 plane i from HBM and
// level l (1..L) computes its plane i-R*l from the 2R+1 planes of level l-1
// around it.  Level L is written to HBM: each cell is read once and written
// once per pass of L steps.  Level l's valid region is the tile less R*l
// cells per side; the last level stores the tile less R*L.  With L = 1 and
// one launch per step this is the per-step method; the persistent method is
// one cooperative launch that runs every pass with a grid barrier in between.
//
// The update sums the shape's groups in the reference gold's order (the shape
// table in perks_shapes.h, including FP32 code's double-literal groups), with
// no FMA contraction.  Domain faces apply the boundary rule at every level.
#pragma once

#include <hip/hip_runtime.h>

#include "grid_sync.h"
#include "j3d_common.h"
#include "perks_shapes.h"
#include "storage_fmt.h"

namespace perks {
namespace j3g {

using fmt::bf16s;
using fmt::f16s;
using fmt::Fmt;

constexpr int kThreads = 256;
constexpr int kRowsPerPass = kThreads / 64;  // tile rows a workgroup covers at once (one wave per row)

struct J3gArgs {
    void* a;                     // field buffers; pass g reads a (g even) or b and writes the other
    void* b;
    int nz, ny, nx;
    int steps;                   // time steps of the whole run
    int levels;                  // L: time steps per full pass (tile geometry uses it)
    int nzs;                     // z-slabs per tile column
    int npass;                   // passes this launch runs
    int step_offset;             // index of this launch's first pass (per-step launches)
    int has_div;
    int sr;                      // stochastic rounding of 16-bit storage
    uint32_t seed;
    GridSync* sync;
    unsigned long long timeout;
};

template <int SHAPE>
constexpr const ShapeDesc& shape3()
{
    return SHAPE == PERKS_SHAPE_3D_STAR_R1     ? kShape3dStarR1
           : SHAPE == PERKS_SHAPE_3D_STAR_R2   ? kShape3dStarR2
           : SHAPE == PERKS_SHAPE_3D_TYPE0_R1  ? kShape3dType0R1
           : SHAPE == PERKS_SHAPE_3D_BOX_R1    ? kShape3dBoxR1
                                               : kShape3dPoissonR1;
}

template <int SHAPE>
constexpr int first_f64_group3()
{
    constexpr const ShapeDesc& s = shape3<SHAPE>();
    for (int g = 0; g < s.ngroups; ++g)
        if (s.g[g].prod_f64) return g;
    return s.ngroups;
}

// One cell.  `get(dz, dy, dx)` returns the neighbour.  Groups before the
// first double-literal group accumulate in C; from that group on (only when C
// is float) the sum is carried in double, as C++ promotion does in the gold.
template <int SHAPE, typename C, typename Get>
__device__ __forceinline__ C eval3(const C (&k)[kMaxGroups], const double (&kd)[kMaxGroups], double kdiv,
                                   bool has_div, Get&& get)
{
    constexpr const ShapeDesc& S = shape3<SHAPE>();
    constexpr int P0 = sizeof(C) == 8 ? S.ngroups : first_f64_group3<SHAPE>();
    C acc = 0;
    double accd = 0.0;
#pragma unroll
    for (int g = 0; g < S.ngroups; ++g) {
        C s = get(S.g[g].off[0][0], S.g[g].off[0][1], S.g[g].off[0][2]);
#pragma unroll
        for (int p = 1; p < S.g[g].npts; ++p) s = s + get(S.g[g].off[p][0], S.g[g].off[p][1], S.g[g].off[p][2]);
        if (g < P0) {
            const C prod = k[g] * s;
            acc = g == 0 ? prod : acc + prod;
        } else if (S.g[g].prod_f64) {
            const double prod = kd[g] * (double)s;
            accd = g == 0 ? prod : (g == P0 ? (double)acc + prod : accd + prod);
        } else {
            const C prod = k[g] * s;
            accd = accd + (double)prod;
        }
    }
    if constexpr (P0 < S.ngroups) return (C)(has_div ? accd / kdiv : accd);
    else return has_div ? acc / (C)kdiv : acc;
}

// Lanes map to x (lane = tile column); the workgroup's waves take rows
// w, w+4, w+8, ...  `lds` holds levels*(2R+1) planes of TY x TX values.
template <int SHAPE, typename S, int BC, int TX, int TY>
__global__ void __launch_bounds__(kThreads) j3g_kernel(J3gArgs a, KCoefs cf)
{
    static_assert(TX == 64, "one lane per tile column");
    using C = typename Fmt<S>::C;
    constexpr const ShapeDesc& SD = shape3<SHAPE>();
    constexpr int R = SD.radius;
    constexpr int W = 2 * R + 1;
    constexpr int NP = TX * TY;
    extern __shared__ double lds_raw[];
    C* ring = reinterpret_cast<C*>(lds_raw);

    C k[kMaxGroups];
    double kd[kMaxGroups];
#pragma unroll
    for (int g = 0; g < kMaxGroups; ++g) {
        k[g] = (C)cf.c[g];
        kd[g] = cf.c[g];
    }
    const double kdiv = cf.divisor;
    const bool hd = a.has_div != 0;

    const int lane = threadIdx.x & 63;
    const int row0 = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int me = xcd_tile_id(blockIdx.x, G);
    const int nz = a.nz, ny = a.ny, nx = a.nx;
    const int L = a.levels;
    const int RL = R * L;
    const int VW = TX - 2 * RL, VH = TY - 2 * RL;
    const int ntx = (nx + VW - 1) / VW, nty = (ny + VH - 1) / VH;
    const int nitems = ntx * nty * a.nzs;

    for (int pass = 0; pass < a.npass; ++pass) {
        const int gp = pass + a.step_offset;
        const int lv = min(L, a.steps - gp * L);
        const S* src = (const S*)((gp & 1) ? a.b : a.a);
        S* dst = (S*)((gp & 1) ? a.a : a.b);
        for (int item = me; item < nitems; item += G) {
            const int tx = item % ntx;
            const int ty = (item / ntx) % nty;
            const int zsl = item / (ntx * nty);
            const int zs = (int)(((long long)zsl * nz) / a.nzs);
            const int ze = (int)(((long long)(zsl + 1) * nz) / a.nzs);
            const int ox = tx * VW - RL, oy = ty * VH - RL;  // domain coordinates of tile cell (0, 0)
            const int dxlo = -ox, dxhi = nx - 1 - ox, dylo = -oy, dyhi = ny - 1 - oy;
            const bool edge = dxlo > 0 || dylo > 0 || dxhi < TX - 1 || dyhi < TY - 1;
            const int x = ox + lane;
            const int xc = min(max(x, 0), nx - 1);
            const int i_first = max(0, zs - R * lv);
            const int i_load_end = min(nz, ze + R * lv);
            const int i_last = ze - 1 + R * lv;

            __syncthreads();  // the previous item may still be reading the rings
            for (int i = i_first; i <= i_last; ++i) {
                if (i < i_load_end) {
                    C* slot = ring + (size_t)(i % W) * NP;
                    const S* plane = src + (size_t)i * ny * nx;
                    for (int ly = row0; ly < TY; ly += kRowsPerPass) {
                        const int y = min(max(oy + ly, 0), ny - 1);
                        slot[ly * TX + lane] = Fmt<S>::ld(plane, (size_t)y * nx + xc);
                    }
                }
                __syncthreads();
                for (int l = 1; l <= lv; ++l) {
                    const int p = i - R * l;
                    const int plo = max(0, zs - R * (lv - l)), phi = min(nz, ze + R * (lv - l));
                    if (p >= plo && p < phi) {
                        const C* in = ring + (size_t)(l - 1) * W * NP;
                        const int m = R * l;  // level l computes the tile less m cells per side
                        const uint32_t key = a.sr ? sr_key((uint32_t)(gp * L + l - 1), a.seed) : 0u;
                        const bool xact = lane >= m && lane < TX - m;
                        for (int ly = m + row0; ly < TY - m; ly += kRowsPerPass) {
                            if (!xact) continue;
                            auto get = [&](int dz, int dy, int dx) -> C {
                                int q = p + dz;
                                int yy = ly + dy, xx = lane + dx;
                                if (BC == PERKS_BC_ZERO) {
                                    if (q < 0 || q >= nz) return (C)0;
                                    if (edge && (yy < dylo || yy > dyhi || xx < dxlo || xx > dxhi)) return (C)0;
                                } else {
                                    q = min(max(q, 0), nz - 1);
                                    if (edge) {
                                        yy = min(max(yy, dylo), dyhi);
                                        xx = min(max(xx, dxlo), dxhi);
                                    }
                                }
                                return in[((size_t)(q % W) * TY + yy) * TX + xx];
                            };
                            C v = eval3<SHAPE, C>(k, kd, kdiv, hd, get);
                            const size_t gidx = ((size_t)p * ny + (size_t)(oy + ly)) * nx + (size_t)x;
                            v = a.sr ? Fmt<S>::qsr(v, (uint32_t)gidx, key) : Fmt<S>::q(v);
                            if (l < lv) {
                                ring[((size_t)l * W + (p % W)) * NP + ly * TX + lane] = v;
                            } else if (p >= zs && p < ze && ly >= RL && ly < TY - RL && lane >= RL &&
                                       lane < TX - RL && oy + ly < ny && x < nx) {
                                Fmt<S>::st(dst, gidx, v);
                            }
                        }
                    }
                    __syncthreads();
                }
            }
        }
        if (pass + 1 < a.npass && !grid_sync(a.sync, G, a.timeout)) return;
    }
}

template <typename S>
struct TileOf3 {
    static constexpr int TX = 64, TY = 32;
};
template <>
struct TileOf3<double> {
    static constexpr int TX = 64, TY = 16;
};

template <int SHAPE, typename S>
void* kptr_bc(int bc)
{
    if (bc == PERKS_BC_ZERO)
        return reinterpret_cast<void*>(&j3g_kernel<SHAPE, S, PERKS_BC_ZERO, TileOf3<S>::TX, TileOf3<S>::TY>);
    return reinterpret_cast<void*>(&j3g_kernel<SHAPE, S, PERKS_BC_CLAMP, TileOf3<S>::TX, TileOf3<S>::TY>);
}

template <typename S>
void* kptr(int shape, int bc)
{
    switch (shape) {
    case PERKS_SHAPE_3D_STAR_R1: return kptr_bc<PERKS_SHAPE_3D_STAR_R1, S>(bc);
    case PERKS_SHAPE_3D_STAR_R2: return kptr_bc<PERKS_SHAPE_3D_STAR_R2, S>(bc);
    case PERKS_SHAPE_3D_TYPE0_R1: return kptr_bc<PERKS_SHAPE_3D_TYPE0_R1, S>(bc);
    case PERKS_SHAPE_3D_BOX_R1: return kptr_bc<PERKS_SHAPE_3D_BOX_R1, S>(bc);
    case PERKS_SHAPE_3D_POISSON_R1: return kptr_bc<PERKS_SHAPE_3D_POISSON_R1, S>(bc);
    default: return nullptr;
    }
}

}  // namespace j3g

// Kernel entry points per storage format (j3d_general_<fmt>.hip); nullptr if not built.
void* j3g_kernel_f64(int shape, int bc);
void* j3g_kernel_f32(int shape, int bc);
void* j3g_kernel_f16(int shape, int bc);
void* j3g_kernel_bf16(int shape, int bc);

}  // namespace perks
