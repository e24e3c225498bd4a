// j2d_stream_kernel.h -- the 2D Jacobi kernel that streams rows through
// registers (tile variant 2; instantiated per storage type in
// j2d_stream_{f64,f32,f16,bf16}.hip).
//
// A work item is a strip of 64 columns (one per lane) and a segment of rows
// of one grid; each wave runs its own items, so there is no LDS and no
// workgroup barrier.  The wave marches down the segment one row per
// iteration: level 0 is the field row read from HBM (prefetched D rows
// ahead), and level l (1..L) computes row i - R*l from the 2R+1 rows of level
// l-1 it keeps in a register ring.  East/west neighbours come from the
// adjacent lanes by DPP whole-wave shifts; box shapes shift every row of
// their window once when it is produced, star shapes shift only the centre
// row.  Each level's valid region shrinks by R columns per side, absorbed by
// a ghost rim of R*L columns, and the segment's rows are extended by R*L ghost
// rows at each end.  Only level L is stored.
//
// Boundaries: at every level, the cells just outside the grid hold what the
// boundary rule reads there -- 0 (zero) or the nearest edge cell (clamp) --
// so the update itself never tests for a face.  Level 0 gets them from the
// load (clamped indices, or 0); higher levels copy the edge row into the rows
// beyond the top and bottom faces when the edge row is produced, and the
// edge column into the lanes beyond the side faces (readlane).  This is the
// per-step kernel's rule applied level by level, so a pass equals L stored
// steps bit for bit: arithmetic is the shape table in order without FMA
// contraction (eval_shape) and every level is rounded to the storage format.
#pragma once

#include <type_traits>

#include "j2d_kernel.h"
#include "perks_device.h"

namespace perks {
namespace j2ds {

using fmt::Fmt;
using j2dk::eval_shape;
using j2dk::shape_desc;

constexpr int kWaves = 4;  // waves per workgroup (each runs its own items)

// sh[R + dx] = the row value at column x + dx, for dx in [-R, R]: lane i
// receives lane i-1 (west) or lane i+1 (east), repeatedly.
template <int R, typename C>
__device__ __forceinline__ void shifts(C c, C (&sh)[2 * R + 1])
{
    sh[R] = c;
#pragma unroll
    for (int k = 1; k <= R; ++k) {
        sh[R - k] = dpp_shift0<kDppWaveShr1>(sh[R - k + 1]);
        sh[R + k] = dpp_shift0<kDppWaveShl1>(sh[R + k - 1]);
    }
}

__device__ __forceinline__ float read_lane(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l)
{
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int SHAPE, typename S, int BC, bool HAS_DIV, int L>
struct Strip {
    using C = typename Fmt<S>::C;
    static constexpr const ShapeDesc& SD = shape_desc<SHAPE>();
    static constexpr int R = SD.radius;
    static constexpr bool BOX = SD.is_box != 0;
    static constexpr int Q = 2 * R + 1;            // rows of a level's window
    static constexpr int D = Q * ((8 + Q - 1) / Q);  // level-0 rows loaded ahead (at least 8, a multiple of Q)
    static constexpr int P = Q + D;                // level-0 ring slots
    static constexpr int UF = P;         // stages per loop trip (a multiple of Q: ring slots are static)
    static constexpr int NS = BOX ? Q : 1;

    const S* src;
    S* dst;
    int nx, ny;
    int x, xc;       // this lane's column, clamped into the grid
    bool xout;       // the column lies outside the grid
    bool xedge;      // the strip reaches past a side face
    int lane_w, lane_e;  // lanes of columns 0 and nx-1 (clamped into the wave)
    bool st_ok;      // this lane stores its column
    int lo[L + 1], hi[L + 1];
    int ys, ye;
    C k[kMaxGroups];
    double kd[kMaxGroups];
    double kdiv;
    bool sr;
    uint32_t key[L + 1];

    C raw[P];         // level-0 rows as loaded
    C win[L][Q][NS];  // level l's window (level 0 only for box shapes); boxes keep the x shifts

    static constexpr int mp(int a) { return ((a % P) + P) % P; }
    static constexpr int mq(int a) { return ((a % Q) + Q) % Q; }

    __device__ __forceinline__ C load(int q) const
    {
        const int qc = q < 0 ? 0 : (q >= ny ? ny - 1 : q);
        const C v = Fmt<S>::ld(src + (size_t)qc * nx, (size_t)xc);
        if constexpr (BC == PERKS_BC_ZERO) return (q < 0 || q >= ny || xout) ? (C)0 : v;
        return v;
    }

    // Level LV's row at relative slot index `rel` (static), column offset dx.
    template <int LV>
    __device__ __forceinline__ C cell(int rel, int dx, const C (&csh)[Q], int centre_rel) const
    {
        if constexpr (BOX) return win[LV][mq(rel)][R + dx];
        else {
            if (rel == centre_rel) return csh[R + dx];
            if constexpr (LV == 0) return raw[mp(rel)];
            else return win[LV][mq(rel)][0];
        }
    }

    // Lanes beyond a side face take the boundary value of this row.
    __device__ __forceinline__ C side_faces(C v) const
    {
        if (!xedge) return v;
        if constexpr (BC == PERKS_BC_ZERO) {
            return xout ? (C)0 : v;
        } else {
            const C vw = read_lane(v, lane_w);
            const C ve = read_lane(v, lane_e);
            return x < 0 ? vw : (x >= nx ? ve : v);
        }
    }

    // Level LV1 = LV + 1 computes row q = i - R*LV1 from level LV's window.
    // Window slots are written every iteration with static indices and
    // selects (no branch around a register-array write): the slot of row q
    // held row q - Q, which no level needs any more, so writing it when q is
    // outside the item's range is harmless.
    template <int U, int LV>
    __device__ __forceinline__ void level(int i)
    {
        if constexpr (LV < L) {
            constexpr int LV1 = LV + 1;
            constexpr int REL = U - R * LV1;  // relative index of row q
            const int q = i - R * LV1;
            C v = (C)0;
            if (q >= lo[LV1] && q < hi[LV1] && q < ny) {
                C csh[Q];
                if constexpr (!BOX) {
                    C c;
                    if constexpr (LV == 0) c = raw[mp(REL)];
                    else c = win[LV][mq(REL)][0];
                    shifts<R>(c, csh);
                }
                v = eval_shape<SHAPE, C, HAS_DIV>(k, kdiv, kd, [&](int dy, int dx) -> C {
                    return cell<LV>(REL + dy, dx, csh, REL);
                });
                if (sr) v = Fmt<S>::qsr(v, (uint32_t)q * (uint32_t)nx + (uint32_t)x, key[LV1]);
                else v = Fmt<S>::q(v);
                if constexpr (LV1 == L) {
                    if (st_ok) Fmt<S>::st(dst + (size_t)q * nx, (size_t)x, v);
                }
            }
            if constexpr (LV1 < L) {
                v = side_faces(v);
                C row[NS];
                if constexpr (BOX) shifts<R>(v, row);
                else row[0] = v;
                // Rows below the bottom face read row ny-1 (clamp) or 0 (zero).
                const int jb = q - (ny - 1);
#pragma unroll
                for (int jj = 1; jj <= R; ++jj)
#pragma unroll
                    for (int t = 0; t < NS; ++t)
                        row[t] = jb == jj ? (BC == PERKS_BC_ZERO ? (C)0 : win[LV1][mq(REL - jj)][t]) : row[t];
#pragma unroll
                for (int t = 0; t < NS; ++t) win[LV1][mq(REL)][t] = row[t];
                // Rows above the top face read row 0 (clamp) or 0 (zero).
#pragma unroll
                for (int j = 1; j <= R; ++j)
#pragma unroll
                    for (int t = 0; t < NS; ++t)
                        win[LV1][mq(REL - j)][t] =
                            q == 0 ? (BC == PERKS_BC_ZERO ? (C)0 : row[t]) : win[LV1][mq(REL - j)][t];
            }
            level<U, LV + 1>(i);
        }
    }

    // Iteration i (stage U = (i - i0) mod UF): level-0 row i has arrived; row
    // i + D is requested into the slot row i - 2R - 1 held.
    template <int U>
    __device__ __forceinline__ void stage(int i)
    {
        raw[mp(U + D)] = load(i + D);
        if constexpr (BOX) {
            C sh[Q];
            shifts<R>(raw[mp(U)], sh);
#pragma unroll
            for (int t = 0; t < Q; ++t) win[0][mq(U)][t] = sh[t];
        }
        level<U, 0>(i);
    }

    template <int U>
    __device__ __forceinline__ bool trip(int i, int iend)
    {
        if constexpr (U < UF) {
            if (i > iend) return false;
            stage<U>(i);
            return trip<U + 1>(i + 1, iend);
        }
        return true;
    }

    // Rows [ys, ye) of the strip whose lane 0 is column xs0; stored columns [vx0, vx1).
    __device__ void run(int lane, int xs0, int vx0, int vx1)
    {
        x = xs0 + lane;
        xc = x < 0 ? 0 : (x >= nx ? nx - 1 : x);
        xout = x < 0 || x >= nx;
        xedge = xs0 < 0 || xs0 + 63 >= nx;
        lane_w = min(max(-xs0, 0), 63);
        lane_e = min(max(nx - 1 - xs0, 0), 63);
        st_ok = x >= vx0 && x < vx1;
#pragma unroll
        for (int l = 0; l <= L; ++l) {
            lo[l] = max(0, ys - R * (L - l));
            // Levels below L also produce the R rows beyond the bottom face.
            hi[l] = l == L ? ye : min(ny + R, ye + R * (L - l));
        }
        const int i0 = ys - R * L;
        const int iend = ye - 1 + R * L;
#pragma unroll
        for (int t = 0; t < D; ++t) raw[t] = load(i0 + t);
        for (int i = i0; trip<0>(i, iend); i += UF) {
        }
    }
};

template <int SHAPE, typename S, int BC, bool HAS_DIV>
__global__ void __launch_bounds__(kWaves * 64) j2d_stream_kernel(J2dArgs a, KCoefs cf)
{
    using C = typename Fmt<S>::C;
    constexpr const ShapeDesc& SD = shape_desc<SHAPE>();
    constexpr int R = SD.radius;
    constexpr int LM = stream_max_levels(R);
    C k[kMaxGroups];
    double kd[kMaxGroups];
#pragma unroll
    for (int g = 0; g < kMaxGroups; ++g) {
        k[g] = (C)cf.c[g];
        kd[g] = cf.c[g];
    }
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int G = gridDim.x;
    const int me = xcd_tile_id(blockIdx.x, G) * kWaves + wave;
    const int W = G * kWaves;

    for (int pass = 0; pass < a.npass; ++pass) {
        const int gp = pass + a.step_offset;
        const int ngrids = a.active[pass];
        const int nitems = a.grids[ngrids - 1].item_end;
        const int RL = R * a.levels;  // strips use the full-pass ghost rim in every pass
        const int VW = 64 - 2 * RL;
        for (int item = me; item < nitems; item += W) {
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
            const int tx = local % ntx, sg = local / ntx;
            const int vx0 = tx * VW, vx1 = min(vx0 + VW, gd.nx);
            const S* src = (const S*)((gp & 1) ? gd.b : gd.a);
            S* dst = (S*)((gp & 1) ? gd.a : gd.b);
            auto go = [&](auto lv) {
                constexpr int LL = decltype(lv)::value;
                Strip<SHAPE, S, BC, HAS_DIV, LL> s;
                s.src = src;
                s.dst = dst;
                s.nx = gd.nx;
                s.ny = gd.ny;
                s.ys = sg * gd.seg;
                s.ye = min(s.ys + gd.seg, gd.ny);
#pragma unroll
                for (int g = 0; g < kMaxGroups; ++g) {
                    s.k[g] = k[g];
                    s.kd[g] = kd[g];
                }
                s.kdiv = cf.divisor;
                s.sr = a.sr != 0;
#pragma unroll
                for (int l = 1; l <= LL; ++l)
                    s.key[l] = a.sr ? sr_key((uint32_t)(gp * a.levels + l - 1), a.seed) : 0u;
                s.run(lane, vx0 - R * LL, vx0, vx1);
            };
            if (L <= 1 || LM < 2) {
                go(std::integral_constant<int, 1>{});
            } else if constexpr (LM >= 2) {
                if (L == 2 || LM < 3) go(std::integral_constant<int, 2>{});
                else if constexpr (LM >= 3) {
                    if (L == 3 || LM < 4) go(std::integral_constant<int, 3>{});
                    else if constexpr (LM >= 4) go(std::integral_constant<int, 4>{});
                }
            }
        }
        if (pass + 1 < a.npass && !grid_sync(a.sync, G, a.timeout)) return;
    }
}

template <int SHAPE, typename S>
void* sptr2(int bc, bool hd)
{
    if (bc == PERKS_BC_ZERO)
        return hd ? reinterpret_cast<void*>(&j2d_stream_kernel<SHAPE, S, PERKS_BC_ZERO, true>)
                  : reinterpret_cast<void*>(&j2d_stream_kernel<SHAPE, S, PERKS_BC_ZERO, false>);
    return hd ? reinterpret_cast<void*>(&j2d_stream_kernel<SHAPE, S, PERKS_BC_CLAMP, true>)
              : reinterpret_cast<void*>(&j2d_stream_kernel<SHAPE, S, PERKS_BC_CLAMP, false>);
}

template <typename S>
void* sptr(int shape, int bc, bool hd)
{
    switch (shape) {
    case PERKS_SHAPE_2D_STAR_R1: return sptr2<PERKS_SHAPE_2D_STAR_R1, S>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R2: return sptr2<PERKS_SHAPE_2D_STAR_R2, S>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R3: return sptr2<PERKS_SHAPE_2D_STAR_R3, S>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R4: return sptr2<PERKS_SHAPE_2D_STAR_R4, S>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R5: return sptr2<PERKS_SHAPE_2D_STAR_R5, S>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R6: return sptr2<PERKS_SHAPE_2D_STAR_R6, S>(bc, hd);
    case PERKS_SHAPE_2D_BOX_R1: return sptr2<PERKS_SHAPE_2D_BOX_R1, S>(bc, hd);
    case PERKS_SHAPE_2D_BOX_R2: return sptr2<PERKS_SHAPE_2D_BOX_R2, S>(bc, hd);
    default: return nullptr;
    }
}

}  // namespace j2ds
}  // namespace perks
