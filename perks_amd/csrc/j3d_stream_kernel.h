// j3d_stream_kernel.h -- the shape-generic 3D Jacobi kernel that streams
// planes through registers (instantiated per storage format in
// j3d_stream_{f64,f32,f16,bf16}.hip; host side in j3d_general.hip).
//
// A work item is an x-y tile of 64 columns (one per lane) by kRows rows (held
// by one wave, one register per row) over a z-slab; each wave runs its own
// items, so there is no LDS and no workgroup barrier.  The wave marches along
// z one plane per iteration: level 0 is the field plane read from HBM
// (requested D planes ahead and kept as loaded until it enters the window,
// where it is widened, so nothing waits on a load before its first use), and
// level l (1..L) computes plane i - R*l from
// the 2R+1 planes of level l-1 it keeps in a register ring.  y neighbours are
// the wave's own rows; x neighbours come from the adjacent lanes by DPP
// whole-wave shifts (a shifted value is formed once per plane and reused by
// every row that needs it).  Level l is valid on the tile less R*l columns and
// rows per side, so a tile stores (64 - 2RL) x (kRows - 2RL) cells; z-slabs
// read R*L ghost planes beyond each end.  Only level L is stored.
//
// Boundaries: at every level the cells just outside the grid hold what the
// boundary rule reads there -- 0 (zero) or the nearest edge cell (clamp) --
// so the update itself never tests for a face.  Level 0 gets them from
// clamped loads (zeroed as the plane enters the window); higher levels copy the edge row, column
// (readlane) or plane into the cells beyond it when it is produced.  The
// update is the shape table in the gold's order without FMA contraction
// (eval3) and every level is rounded to the storage format, so a pass equals
// L stored steps of the per-step method bit for bit.
#pragma once

#include <type_traits>

#include "j3d_general_kernel.h"
#include "perks_device.h"

namespace perks {
namespace j3s {

using fmt::Fmt;
using j3g::eval3;
using j3g::J3gArgs;
using j3g::shape3;

constexpr int kWaves = 4;  // waves per workgroup (each runs its own items)
constexpr int kRows = 8;   // tile rows per wave (the short tile; kRowsTall for the tall one)
constexpr int kRowsTall = 16;

// Deepest pass the kernel is built for, by radius (registers).
__host__ __device__ constexpr int stream3_max_levels(int radius) { return radius == 1 ? 2 : 1; }

// v moved by dx columns: lane i receives lane i+dx (dx static after unrolling).
template <int R, typename C>
__device__ __forceinline__ C xs(C v, int dx)
{
    C w = v;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (dx < 0 && k < -dx) w = dpp_shift0<kDppWaveShr1>(w);
        if (dx > 0 && k < dx) w = dpp_shift0<kDppWaveShl1>(w);
    }
    return w;
}

__device__ __forceinline__ float rd_lane(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rd_lane(double v, int l)
{
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int SHAPE, typename S, int BC, int L, int NY>
struct Tile3 {
    using C = typename Fmt<S>::C;
    static constexpr const ShapeDesc& SD = shape3<SHAPE>();
    static constexpr int R = SD.radius;
    static constexpr int Q = 2 * R + 1;  // planes of a level's window
    static constexpr int D = Q;          // level-0 planes loaded ahead
    static constexpr int UF = D;         // stages per loop trip (a multiple of Q: ring slots are static)

    const S* src;
    S* dst;
    int nz, ny, nx;
    size_t plane;
    int x;              // this lane's column
    bool xout, xedge;   // column outside the grid; tile reaches past a side face
    int lane_w, lane_e; // lanes of columns 0 and nx-1 (clamped into the wave)
    int ys0;            // grid row of tile row 0
    bool yedge;         // tile reaches past a top/bottom face
    uint32_t roff[NY];  // element offset of (clamped row, clamped column) in a plane
    bool rin[NY];       // row inside the grid
    bool st_x;          // this lane stores
    int st_r0, st_r1;   // tile rows stored: [st_r0, st_r1)
    int lo[L + 1], hi[L + 1];
    C k[kMaxGroups];
    double kd[kMaxGroups];
    double kdiv;
    bool hd, sr;
    uint32_t key[L + 1];

    using Raw = typename Fmt<S>::Raw;
    Raw fl[D][NY];    // level-0 planes in flight, as loaded (plane i in slot (i - i0) mod D)
    C win[L][Q][NY];  // level l's window

    static constexpr int mq(int a) { return ((a % Q) + Q) % Q; }

    __device__ __forceinline__ void load(int q, Raw (&dstp)[NY]) const
    {
        const int qc = q < 0 ? 0 : (q >= nz ? nz - 1 : q);
        const S* base = src + (size_t)qc * plane;
#pragma unroll
        for (int r = 0; r < NY; ++r) dstp[r] = Fmt<S>::ld_raw(base, (size_t)roff[r]);
    }

    // Plane q of level 0 enters the window at its first use, D planes after
    // its load was issued: widened, and 0 outside the grid (zero boundary).
    __device__ __forceinline__ void arrive(int q, const Raw (&w)[NY], C (&out)[NY]) const
    {
        const bool qin = q >= 0 && q < nz;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const C v = Fmt<S>::widen(w[r]);
            if constexpr (BC == PERKS_BC_ZERO) out[r] = (qin && rin[r] && !xout) ? v : (C)0;
            else out[r] = v;
        }
    }

    template <int LV>
    __device__ __forceinline__ C at(int rel, int r) const
    {
        return win[LV][mq(rel)][r];
    }

    // Level LV1 = LV + 1 computes plane q = i - R*LV1 from level LV's window.
    // Window slots are written every iteration with static indices and
    // selects: the slot of plane q held plane q - Q, which no level needs any
    // more, so writing it when q is outside the item's range is harmless.
    template <int U, int LV>
    __device__ __forceinline__ void level(int i)
    {
        if constexpr (LV < L) {
            constexpr int LV1 = LV + 1;
            constexpr int REL = U - R * LV1;  // relative index of plane q
            constexpr int RA = R * LV1, RB = NY - R * LV1;  // rows this level computes
            const int q = i - R * LV1;
            C out[NY];
#pragma unroll
            for (int r = 0; r < NY; ++r) out[r] = (C)0;
            if (q >= lo[LV1] && q < hi[LV1] && q < nz) {
#pragma unroll
                for (int r = RA; r < RB; ++r) {
                    C v = eval3<SHAPE, C>(k, kd, kdiv, hd, [&](int dz, int dy, int dx) -> C {
                        return xs<R>(at<LV>(REL + dz, r + dy), dx);
                    });
                    if (sr)
                        v = Fmt<S>::qsr(v, (uint32_t)(((size_t)q * ny + (size_t)(ys0 + r)) * nx + (size_t)x),
                                        key[LV1]);
                    else v = Fmt<S>::q(v);
                    out[r] = v;
                }
                if constexpr (LV1 == L) {
#pragma unroll
                    for (int r = RA; r < RB; ++r)
                        if (r >= st_r0 && r < st_r1 && st_x)
                            Fmt<S>::st(dst, ((size_t)q * ny + (size_t)(ys0 + r)) * nx + (size_t)x, out[r]);
                }
            }
            if constexpr (LV1 < L) {
                // Rows beyond the top/bottom faces take row 0 / ny-1 (clamp) or 0.
                if (yedge) {
                    const int r0 = -ys0, r1 = ny - 1 - ys0;
                    C e0 = (C)0, e1 = (C)0;
                    if constexpr (BC == PERKS_BC_CLAMP) {
#pragma unroll
                        for (int r = RA; r < RB; ++r) {
                            e0 = r == r0 ? out[r] : e0;
                            e1 = r == r1 ? out[r] : e1;
                        }
                    }
#pragma unroll
                    for (int r = RA; r < RB; ++r) out[r] = r < r0 ? e0 : (r > r1 ? e1 : out[r]);
                }
                // Columns beyond the side faces take column 0 / nx-1 (clamp) or 0.
                if (xedge) {
#pragma unroll
                    for (int r = RA; r < RB; ++r) {
                        if constexpr (BC == PERKS_BC_ZERO) {
                            out[r] = xout ? (C)0 : out[r];
                        } else {
                            const C vw = rd_lane(out[r], lane_w);
                            const C ve = rd_lane(out[r], lane_e);
                            out[r] = x < 0 ? vw : (x >= nx ? ve : out[r]);
                        }
                    }
                }
                // Planes beyond the bottom face read plane nz-1 (clamp) or 0.
                const int jb = q - (nz - 1);
#pragma unroll
                for (int jj = 1; jj <= R; ++jj)
#pragma unroll
                    for (int r = RA; r < RB; ++r)
                        out[r] = jb == jj ? (BC == PERKS_BC_ZERO ? (C)0 : win[LV1][mq(REL - jj)][r]) : out[r];
#pragma unroll
                for (int r = 0; r < NY; ++r) win[LV1][mq(REL)][r] = out[r];
                // Planes above the top face read plane 0 (clamp) or 0.
#pragma unroll
                for (int j = 1; j <= R; ++j)
#pragma unroll
                    for (int r = 0; r < NY; ++r)
                        win[LV1][mq(REL - j)][r] =
                            q == 0 ? (BC == PERKS_BC_ZERO ? (C)0 : out[r]) : win[LV1][mq(REL - j)][r];
            }
            level<U, LV + 1>(i);
        }
    }

    // Iteration i (stage U = (i - i0) mod UF): level-0 plane i, loaded D
    // iterations ago, enters the window, and plane i + D is requested into
    // its in-flight slot.
    template <int U>
    __device__ __forceinline__ void stage(int i)
    {
        // The next load is issued before plane i is waited for, so D planes
        // stay in flight across the wait.
        Raw w[NY];
#pragma unroll
        for (int r = 0; r < NY; ++r) w[r] = fl[U][r];
        load(i + D, fl[U]);
        arrive(i, w, win[0][mq(U)]);
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

    // Tile whose lane 0 is column xs0 and row 0 is grid row ys0; stores
    // columns [vx0, vx1) x rows [vy0, vy1) of planes [zs, ze).
    __device__ void run(int lane, int xs0, int vx0, int vx1, int vy0, int vy1, int zs, int ze)
    {
        x = xs0 + lane;
        const int xc = x < 0 ? 0 : (x >= nx ? nx - 1 : x);
        xout = x < 0 || x >= nx;
        xedge = xs0 < 0 || xs0 + 63 >= nx;
        lane_w = min(max(-xs0, 0), 63);
        lane_e = min(max(nx - 1 - xs0, 0), 63);
        yedge = ys0 < 0 || ys0 + NY - 1 >= ny;
        st_x = x >= vx0 && x < vx1;
        st_r0 = vy0 - ys0;
        st_r1 = vy1 - ys0;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int y = ys0 + r;
            const int yc = y < 0 ? 0 : (y >= ny ? ny - 1 : y);
            rin[r] = y >= 0 && y < ny;
            roff[r] = (uint32_t)(yc * nx + xc);
        }
#pragma unroll
        for (int l = 0; l <= L; ++l) {
            lo[l] = max(0, zs - R * (L - l));
            hi[l] = l == L ? ze : min(nz, ze + R * (L - l));
        }
        const int i0 = zs - R * L;
        const int iend = ze - 1 + R * L;
#pragma unroll
        for (int t = 0; t < D; ++t) load(i0 + t, fl[t]);
        for (int i = i0; trip<0>(i, iend); i += UF) {
        }
    }
};

template <int SHAPE, typename S, int BC, int NY>
__global__ void __launch_bounds__(kWaves * 64) j3s_kernel(J3gArgs a, KCoefs cf)
{
    using C = typename Fmt<S>::C;
    constexpr int R = shape3<SHAPE>().radius;
    constexpr int LM = stream3_max_levels(R);
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
    const int nz = a.nz, ny = a.ny, nx = a.nx;
    const int RL = R * a.levels;  // tiles use the full-pass ghost rim in every pass
    const int VW = 64 - 2 * RL, VH = NY - 2 * RL;
    const int ntx = (nx + VW - 1) / VW, nty = (ny + VH - 1) / VH;
    const int nitems = ntx * nty * a.nzs;

    for (int pass = 0; pass < a.npass; ++pass) {
        const int gp = pass + a.step_offset;
        const int L = min(a.levels, a.steps - gp * a.levels);
        const S* src = (const S*)((gp & 1) ? a.b : a.a);
        S* dst = (S*)((gp & 1) ? a.a : a.b);
        for (int item = me; item < nitems; item += W) {
            const int tx = item % ntx;
            const int ty = (item / ntx) % nty;
            const int zsl = item / (ntx * nty);
            const int zs = (int)(((long long)zsl * nz) / a.nzs);
            const int ze = (int)(((long long)(zsl + 1) * nz) / a.nzs);
            const int vx0 = tx * VW, vx1 = min(vx0 + VW, nx);
            const int vy0 = ty * VH, vy1 = min(vy0 + VH, ny);
            auto go = [&](auto lv) {
                constexpr int LL = decltype(lv)::value;
                Tile3<SHAPE, S, BC, LL, NY> t;
                t.src = src;
                t.dst = dst;
                t.nz = nz;
                t.ny = ny;
                t.nx = nx;
                t.plane = (size_t)ny * nx;
                t.ys0 = vy0 - R * LL;
#pragma unroll
                for (int g = 0; g < kMaxGroups; ++g) {
                    t.k[g] = k[g];
                    t.kd[g] = kd[g];
                }
                t.kdiv = cf.divisor;
                t.hd = a.has_div != 0;
                t.sr = a.sr != 0;
#pragma unroll
                for (int l = 1; l <= LL; ++l)
                    t.key[l] = a.sr ? sr_key((uint32_t)(gp * a.levels + l - 1), a.seed) : 0u;
                t.run(lane, vx0 - R * LL, vx0, vx1, vy0, vy1, zs, ze);
            };
            if (L <= 1 || LM < 2) go(std::integral_constant<int, 1>{});
            else if constexpr (LM >= 2) go(std::integral_constant<int, 2>{});
        }
        if (pass + 1 < a.npass && !grid_sync(a.sync, G, a.timeout)) return;
    }
}

template <int SHAPE, typename S, int NY>
void* sptr_bc(int bc)
{
    if (bc == PERKS_BC_ZERO) return reinterpret_cast<void*>(&j3s_kernel<SHAPE, S, PERKS_BC_ZERO, NY>);
    return reinterpret_cast<void*>(&j3s_kernel<SHAPE, S, PERKS_BC_CLAMP, NY>);
}

// The tall tile is built for FP32 arithmetic (FP32, FP16, BF16 storage):
// in FP64 its windows take every VGPR of a wave and some shapes spill.
template <int SHAPE, typename S>
void* sptr_rows(int bc, int rows)
{
    if constexpr (std::is_same<S, double>::value) {
        return rows == kRowsTall ? nullptr : sptr_bc<SHAPE, S, kRows>(bc);
    } else {
        return rows == kRowsTall ? sptr_bc<SHAPE, S, kRowsTall>(bc) : sptr_bc<SHAPE, S, kRows>(bc);
    }
}

// rows: kRows or kRowsTall tile rows per wave (nullptr: not built).
template <typename S>
void* sptr(int shape, int bc, int rows)
{
    switch (shape) {
    case PERKS_SHAPE_3D_STAR_R1: return sptr_rows<PERKS_SHAPE_3D_STAR_R1, S>(bc, rows);
    case PERKS_SHAPE_3D_STAR_R2: return sptr_rows<PERKS_SHAPE_3D_STAR_R2, S>(bc, rows);
    case PERKS_SHAPE_3D_TYPE0_R1: return sptr_rows<PERKS_SHAPE_3D_TYPE0_R1, S>(bc, rows);
    case PERKS_SHAPE_3D_BOX_R1: return sptr_rows<PERKS_SHAPE_3D_BOX_R1, S>(bc, rows);
    case PERKS_SHAPE_3D_POISSON_R1: return sptr_rows<PERKS_SHAPE_3D_POISSON_R1, S>(bc, rows);
    default: return nullptr;
    }
}

}  // namespace j3s

// Kernel entry points per storage format (j3d_stream_<fmt>.hip).
void* j3s_kernel_f64(int shape, int bc, int rows);
void* j3s_kernel_f32(int shape, int bc, int rows);
void* j3s_kernel_f16(int shape, int bc, int rows);
void* j3s_kernel_bf16(int shape, int bc, int rows);

}  // namespace perks
