// j2d_kernel.h -- the 2D Jacobi kernel template (instantiated per storage type
// in j2d_f64.hip, j2d_f32.hip, j2d_f16.hip, j2d_bf16.hip so they build in parallel).
//
// A workgroup stages a TX x TY tile of the field, including a ghost rim of
// R*L cells (R = shape radius, L = time steps in the pass), in LDS, runs the L
// steps in LDS (ping-pong between two LDS tiles, one workgroup barrier per
// step), and writes back the tile's interior.  Arithmetic follows the shape
// table in order (no FMA contraction); FP16/BF16 storage is widened to FP32
// for the update and rounded to nearest-even on store.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "j2d.h"
#include "storage_fmt.h"

namespace perks {
namespace j2dk {

using fmt::bf16s;
using fmt::f16s;
using fmt::Fmt;

// ---- one cell update from the shape table ------------------------------------
template <int SHAPE>
constexpr const ShapeDesc& shape_desc()
{
    return SHAPE == PERKS_SHAPE_2D_STAR_R1   ? kShape2dStarR1
           : SHAPE == PERKS_SHAPE_2D_STAR_R2 ? kShape2dStarR2
           : SHAPE == PERKS_SHAPE_2D_STAR_R3 ? kShape2dStarR3
           : SHAPE == PERKS_SHAPE_2D_STAR_R4 ? kShape2dStarR4
           : SHAPE == PERKS_SHAPE_2D_STAR_R5 ? kShape2dStarR5
           : SHAPE == PERKS_SHAPE_2D_STAR_R6 ? kShape2dStarR6
           : SHAPE == PERKS_SHAPE_2D_BOX_R1  ? kShape2dBoxR1
                                             : kShape2dBoxR2;
}

template <int SHAPE>
constexpr int first_f64_group()
{
    constexpr const ShapeDesc& s = shape_desc<SHAPE>();
    for (int g = 0; g < s.ngroups; ++g)
        if (s.g[g].prod_f64) return g;
    return s.ngroups;
}

// Evaluates the shape at one cell.  `get(dy, dx)` returns the neighbour.
// Groups before the first double-literal group accumulate in C; from that
// group on (only possible when C is float) the sum is carried in double.
template <int SHAPE, typename C, bool HAS_DIV, typename Get>
__device__ __forceinline__ C eval_shape(const C (&k)[kMaxGroups], double kd_div, const double (&kd)[kMaxGroups],
                                        Get&& get)
{
    constexpr const ShapeDesc& S = shape_desc<SHAPE>();
    constexpr int P0 = sizeof(C) == 8 ? S.ngroups : first_f64_group<SHAPE>();
    C acc = 0;
    double accd = 0.0;
#pragma unroll
    for (int g = 0; g < S.ngroups; ++g) {
        C s = get(S.g[g].off[0][1], S.g[g].off[0][2]);
#pragma unroll
        for (int p = 1; p < S.g[g].npts; ++p) s = s + get(S.g[g].off[p][1], S.g[g].off[p][2]);
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
    if (P0 < S.ngroups) return (C)(HAS_DIV ? accd / kd_div : accd);
    return HAS_DIV ? acc / (C)kd_div : acc;
}

// ---- the kernel ---------------------------------------------------------------
template <int SHAPE, typename S, int BC, int TX, int TY, bool HAS_DIV>
__global__ void __launch_bounds__(256) j2d_kernel(J2dArgs a, KCoefs cf)
{
    using C = typename Fmt<S>::C;
    constexpr const ShapeDesc& SD = shape_desc<SHAPE>();
    constexpr int R = SD.radius;
    __shared__ C tile[2][TY][TX];

    C k[kMaxGroups];
    double kd[kMaxGroups];
#pragma unroll
    for (int g = 0; g < kMaxGroups; ++g) {
        k[g] = (C)cf.c[g];
        kd[g] = cf.c[g];
    }
    const double kdiv = cf.divisor;

    // Lanes map to tile columns (lane, lane+64, ...), the four waves to rows
    // w, w+4, w+8, ...: no per-cell index division.
    constexpr int CX = TX / 64;
    constexpr int kRowStep = 4;
    static_assert(TX % 64 == 0, "tile width is a multiple of the wave");
    const int lane = threadIdx.x & 63;
    const int row0 = threadIdx.x >> 6;
    const int G = gridDim.x;
    const int me = xcd_tile_id(blockIdx.x, G);

    for (int pass = 0; pass < a.npass; ++pass) {
        const int gp = pass + a.step_offset;  // global pass index (per-step launches pass one each)
        const int ngrids = a.active[pass];
        const int nitems = a.grids[ngrids - 1].item_end;
        // Tile geometry always uses the full pass depth, so the tiling (and the
        // host's item count) is the same in every pass; a shorter last pass
        // simply runs fewer levels inside the same ghost rim.
        const int RL = R * a.levels;
        const int VW = TX - 2 * RL, VH = TY - 2 * RL;
        for (int item = me; item < nitems; item += G) {
            int lo = 0, hi = ngrids - 1;  // first grid whose item_end > item
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (a.grids[mid].item_end > item) hi = mid;
                else lo = mid + 1;
            }
            const J2dGrid gd = a.grids[lo];
            const int L = min(a.levels, gd.steps - gp * a.levels);
            const int local = item - (gd.item_end - gd.tiles);
            const int ntx = (gd.nx + VW - 1) / VW;
            const int tx = local % ntx, ty = local / ntx;
            const int vx0 = tx * VW, vy0 = ty * VH;
            const int ox = vx0 - RL, oy = vy0 - RL;  // domain coordinates of tile cell (0,0)
            const S* src = (const S*)((gp & 1) ? gd.b : gd.a);
            S* dst = (S*)((gp & 1) ? gd.a : gd.b);
            const int nx = gd.nx, ny = gd.ny;

            __syncthreads();  // the previous item may still be reading the tile
            for (int ly = row0; ly < TY; ly += kRowStep) {
                const S* rowp = src + (size_t)min(max(oy + ly, 0), ny - 1) * nx;
#pragma unroll
                for (int cx = 0; cx < CX; ++cx) {
                    const int lx = lane + 64 * cx;
                    tile[0][ly][lx] = Fmt<S>::ld(rowp, (size_t)min(max(ox + lx, 0), nx - 1));
                }
            }
            __syncthreads();
            // Domain edges in tile coordinates; tiles that reach past an edge
            // clamp (or zero) the neighbour reads, interior tiles read directly.
            const int dxlo = -ox, dxhi = nx - 1 - ox, dylo = -oy, dyhi = ny - 1 - oy;
            const bool edge = dxlo > 0 || dylo > 0 || dxhi < TX - 1 || dyhi < TY - 1;
            for (int l = 0; l < L; ++l) {
                const int cur = l & 1;
                // stochastic rounding key of this level's time step (uniform)
                const uint32_t key = a.sr ? sr_key((uint32_t)(gp * a.levels + l), a.seed) : 0u;
                const int m = R * (l + 1);  // level l+1 is valid on the tile less m cells per side
                for (int ly = m + row0; ly < TY - m; ly += kRowStep)
#pragma unroll
                for (int cx = 0; cx < CX; ++cx) {
                    const int lx = lane + 64 * cx;
                    if (lx < m || lx >= TX - m) continue;
                    C v;
                    if (!edge) {
                        v = eval_shape<SHAPE, C, HAS_DIV>(k, kdiv, kd, [&](int dy, int dx) {
                            return tile[cur][ly + dy][lx + dx];
                        });
                    } else {
                        v = eval_shape<SHAPE, C, HAS_DIV>(k, kdiv, kd, [&](int dy, int dx) {
                            int yy = ly + dy, xx = lx + dx;
                            if (BC == PERKS_BC_ZERO) {
                                const bool out = yy < dylo || yy > dyhi || xx < dxlo || xx > dxhi;
                                return out ? (C)0 : tile[cur][yy][xx];
                            }
                            yy = min(max(yy, dylo), dyhi);
                            xx = min(max(xx, dxlo), dxhi);
                            return tile[cur][yy][xx];
                        });
                    }
                    // Each step is a stored field: round to the storage format.
                    if (a.sr) {
                        const uint32_t idx = (uint32_t)(oy + ly) * (uint32_t)nx + (uint32_t)(ox + lx);
                        tile[cur ^ 1][ly][lx] = Fmt<S>::qsr(v, idx, key);
                    } else {
                        tile[cur ^ 1][ly][lx] = Fmt<S>::q(v);
                    }
                }
                __syncthreads();
            }
            const int fin = L & 1;
            for (int ly = RL + row0; ly < TY - RL && oy + ly < ny; ly += kRowStep) {
                S* rowp = dst + (size_t)(oy + ly) * nx;
#pragma unroll
                for (int cx = 0; cx < CX; ++cx) {
                    const int lx = lane + 64 * cx;
                    if (lx >= RL && lx < TX - RL && ox + lx < nx) Fmt<S>::st(rowp, (size_t)(ox + lx), tile[fin][ly][lx]);
                }
            }
        }
        if (pass + 1 < a.npass && !grid_sync(a.sync, G, a.timeout)) return;
    }
}

// Tile shapes (TX x TY, two LDS tiles of them): variant 0 (128 x 32; 64 x 32
// for FP64: four workgroups' LDS per CU, which hides latency best in
// measurements at 8192^2) and variant 1, twice the height (less ghost-rim
// work per cell, two workgroups per CU).
template <typename S, int V>
struct TileOf {
    static constexpr int TX = 128, TY = V == 0 ? 32 : 64;
};
template <int V>
struct TileOf<double, V> {
    static constexpr int TX = 64, TY = V == 0 ? 32 : 64;
};

template <int SHAPE, typename S, int BC, bool HD, int V>
void* kptr3()
{
    return reinterpret_cast<void*>(&j2d_kernel<SHAPE, S, BC, TileOf<S, V>::TX, TileOf<S, V>::TY, HD>);
}

template <int SHAPE, typename S, int V>
void* kptr2(int bc, bool hd)
{
    if (bc == PERKS_BC_ZERO)
        return hd ? kptr3<SHAPE, S, PERKS_BC_ZERO, true, V>() : kptr3<SHAPE, S, PERKS_BC_ZERO, false, V>();
    return hd ? kptr3<SHAPE, S, PERKS_BC_CLAMP, true, V>() : kptr3<SHAPE, S, PERKS_BC_CLAMP, false, V>();
}

template <typename S, int V>
void* kptr_v(int shape, int bc, bool hd)
{
    switch (shape) {
    case PERKS_SHAPE_2D_STAR_R1: return kptr2<PERKS_SHAPE_2D_STAR_R1, S, V>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R2: return kptr2<PERKS_SHAPE_2D_STAR_R2, S, V>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R3: return kptr2<PERKS_SHAPE_2D_STAR_R3, S, V>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R4: return kptr2<PERKS_SHAPE_2D_STAR_R4, S, V>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R5: return kptr2<PERKS_SHAPE_2D_STAR_R5, S, V>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R6: return kptr2<PERKS_SHAPE_2D_STAR_R6, S, V>(bc, hd);
    case PERKS_SHAPE_2D_BOX_R1: return kptr2<PERKS_SHAPE_2D_BOX_R1, S, V>(bc, hd);
    case PERKS_SHAPE_2D_BOX_R2: return kptr2<PERKS_SHAPE_2D_BOX_R2, S, V>(bc, hd);
    default: return nullptr;
    }
}

template <typename S>
void* kptr(int shape, int bc, bool hd, int tile)
{
    return tile == 1 ? kptr_v<S, 1>(shape, bc, hd) : kptr_v<S, 0>(shape, bc, hd);
}

}  // namespace j2dk
}  // namespace perks
