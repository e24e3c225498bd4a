// j2d_stream_kernel.h -- the 2D Jacobi kernel that streams rows through
// registers (tile variant 2; instantiated per storage type in
// j2d_stream_{f64,f32,f16,bf16}.hip).
//
// A work item is a strip of 64 columns (one per lane) and a segment of rows
// of one grid; each wave runs its own items, so there is no LDS and no
// workgroup barrier.  The wave marches down the segment one row per
// iteration: level 0 is the field row read from HBM (requested D rows ahead
// and kept as loaded until it enters the window, where it is widened, so
// nothing waits on a load before its first use), and level l (1..L) computes
// row i - R*l from the 2R+1 rows of level
// l-1 it keeps in a register ring.  East/west neighbours come from the
// adjacent lanes by DPP whole-wave shifts; box shapes shift every row of
// their window once when it is produced, star shapes shift only the centre
// row.  Each level's valid region shrinks by R columns per side, absorbed by
// a ghost rim of R*L columns, and the segment's rows are extended by R*L ghost
// rows at each end.  Only level L is stored.
//
// Boundaries: at every level, the cells just outside the grid hold what the
// boundary rule reads there -- 0 (zero) or the nearest edge cell (clamp) --
// so the update itself never tests for a face.  Level 0 gets them from
// clamped loads (zeroed as the row enters the window); higher levels copy the edge row into the rows
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
    static constexpr int UF = D;  // stages per loop trip: a multiple of Q and D, so every ring slot is static
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

    using Raw = typename Fmt<S>::Raw;
    Raw fl[D];        // level-0 rows in flight, as loaded (row i in slot (i - i0) mod D)
    C win[L][Q][NS];  // level l's window; boxes keep the x shifts

    static constexpr int mq(int a) { return ((a % Q) + Q) % Q; }

    __device__ __forceinline__ Raw load(int q) const
    {
        const int qc = q < 0 ? 0 : (q >= ny ? ny - 1 : q);
        return Fmt<S>::ld_raw(src + (size_t)qc * nx, (size_t)xc);
    }

    // Row q of level 0 from its loaded word, at its first use (D rows after
    // its load was issued): widened, and 0 outside the grid (zero boundary).
    __device__ __forceinline__ C arrive(int q, Raw w) const
    {
        const C v = Fmt<S>::widen(w);
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
            return win[LV][mq(rel)][0];
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
                    shifts<R>(win[LV][mq(REL)][0], csh);
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

    // Iteration i (stage U = (i - i0) mod UF): level-0 row i, loaded D
    // iterations ago, enters the window, and row i + D is requested into its
    // in-flight slot.
    template <int U>
    __device__ __forceinline__ void stage(int i)
    {
        // The next load is issued before row i is waited for, so D loads
        // stay in flight across the wait.
        const Raw w = fl[U];
        fl[U] = load(i + D);
        const C c = arrive(i, w);
        if constexpr (BOX) {
            C sh[Q];
            shifts<R>(c, sh);
#pragma unroll
            for (int t = 0; t < Q; ++t) win[0][mq(U)][t] = sh[t];
        } else {
            win[0][mq(U)][0] = c;
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
        for (int t = 0; t < D; ++t) fl[t] = load(i0 + t);
        for (int i = i0; trip<0>(i, iend); i += UF) {
        }
    }
};

// ---- two columns per lane ---------------------------------------------------------
// Storage access to the cells of columns x0 (even) and x0 + 1 in one load or
// store, and their rounding to the storage format.
template <typename S>
struct Pair2;
template <>
struct Pair2<float> {
    using Raw = float2;
    __device__ static Raw ld_raw(const float* row, size_t x0) { return *reinterpret_cast<const float2*>(row + x0); }
    __device__ static void unpack(Raw w, float& a, float& b)
    {
        a = w.x;
        b = w.y;
    }
    __device__ static void ld(const float* row, size_t x0, float& a, float& b)
    {
        const float2 v = *reinterpret_cast<const float2*>(row + x0);
        a = v.x;
        b = v.y;
    }
    __device__ static void st(float* row, size_t x0, float a, float b)
    {
        *reinterpret_cast<float2*>(row + x0) = make_float2(a, b);
    }
    __device__ static void q2(float&, float&, bool, uint32_t, uint32_t) {}
};
template <typename S16, int FMT>
struct Pair16 {
    using Raw = uint32_t;
    __device__ static Raw ld_raw(const S16* row, size_t x0) { return *reinterpret_cast<const uint32_t*>(row + x0); }
    __device__ static void unpack(Raw w, float& a, float& b)
    {
        a = widen_lo<FMT>(w);
        b = widen_hi<FMT>(w);
    }
    __device__ static void ld(const S16* row, size_t x0, float& a, float& b)
    {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(row + x0);
        a = widen_lo<FMT>(w);
        b = widen_hi<FMT>(w);
    }
    // a and b are already storage values: the conversion is exact.
    __device__ static void st(S16* row, size_t x0, float a, float b)
    {
        *reinterpret_cast<uint32_t*>(row + x0) = Store16<FMT>::narrow2(a, b);
    }
    __device__ static void q2(float& a, float& b, bool sr, uint32_t idx0, uint32_t key)
    {
        if (sr) {
            a = Fmt<S16>::qsr(a, idx0, key);
            b = Fmt<S16>::qsr(b, idx0 + 1u, key);
        } else {
            const uint32_t w = Store16<FMT>::narrow2(a, b);
            a = widen_lo<FMT>(w);
            b = widen_hi<FMT>(w);
        }
    }
};
template <>
struct Pair2<fmt::f16s> : Pair16<fmt::f16s, PERKS_F16> {};
template <>
struct Pair2<fmt::bf16s> : Pair16<fmt::bf16s, PERKS_BF16> {};

// v moved by s lanes (lane i receives lane i + s), s static after unrolling.
template <int R, typename C>
__device__ __forceinline__ C lane_shift(C v, int s)
{
    C w = v;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (s < 0 && k < -s) w = dpp_shift0<kDppWaveShr1>(w);
        if (s > 0 && k < s) w = dpp_shift0<kDppWaveShl1>(w);
    }
    return w;
}

// The pair (columns x0 + dx, x1 + dx) for dx in [-R, R] from the pair (a, b)
// of columns (x0, x1): column x0 + m is lane floor(m/2), half m mod 2.
template <int R, typename C>
__device__ __forceinline__ void pair_shifts(C a, C b, C (&sa)[2 * R + 1], C (&sb)[2 * R + 1])
{
#pragma unroll
    for (int dx = -R; dx <= R; ++dx) {
        const int m0 = dx, m1 = dx + 1;
        const int s0 = m0 >= 0 ? m0 / 2 : -((1 - m0) / 2);
        const int s1 = m1 >= 0 ? m1 / 2 : -((1 - m1) / 2);
        sa[R + dx] = lane_shift<R, C>((m0 - 2 * s0) ? b : a, s0);
        sb[R + dx] = lane_shift<R, C>((m1 - 2 * s1) ? b : a, s1);
    }
}

// As Strip, with each lane holding two neighbouring columns (a: even x0,
// b: x0 + 1): one load, store and set of lane shifts serves two cells.  The
// strip is 128 columns wide and starts on an even column; the grid width must
// be even.
template <int SHAPE, typename S, int BC, bool HAS_DIV, int L>
struct Strip2 {
    using C = typename Fmt<S>::C;
    static constexpr const ShapeDesc& SD = shape_desc<SHAPE>();
    static constexpr int R = SD.radius;
    static constexpr bool BOX = SD.is_box != 0;
    static constexpr int Q = 2 * R + 1;
    static constexpr int D = Q * ((8 + Q - 1) / Q);
    static constexpr int UF = D;
    static constexpr int NS = BOX ? Q : 1;

    const S* src;
    S* dst;
    int nx, ny;
    int x0;              // this lane's even column
    bool pout, xedge;    // pair outside the grid; strip reaches past a side face
    int lane_w, lane_e;  // lanes holding columns 0 (half a) and nx-1 (half b)
    bool st_ok;
    int lo[L + 1], hi[L + 1];
    int ys, ye;
    C k[kMaxGroups];
    double kd[kMaxGroups];
    double kdiv;
    bool sr;
    uint32_t key[L + 1];

    using Raw = typename Pair2<S>::Raw;
    Raw fl[D];  // level-0 rows in flight, as loaded (row i in slot (i - i0) mod D)
    C win[L][Q][NS][2];

    static constexpr int mq(int v) { return ((v % Q) + Q) % Q; }

    __device__ __forceinline__ Raw load(int q) const
    {
        const int qc = q < 0 ? 0 : (q >= ny ? ny - 1 : q);
        const int xc = x0 < 0 ? 0 : (x0 > nx - 2 ? nx - 2 : x0);
        return Pair2<S>::ld_raw(src + (size_t)qc * nx, (size_t)xc);
    }

    // Row q of level 0 from its loaded word, at its first use (D rows after
    // its load was issued): unpacked, with the boundary rule applied to pairs
    // outside the grid.
    __device__ __forceinline__ void arrive(int q, Raw w, C& oa, C& ob) const
    {
        C a, b;
        Pair2<S>::unpack(w, a, b);
        if constexpr (BC == PERKS_BC_ZERO) {
            const bool z = q < 0 || q >= ny || pout;
            oa = z ? (C)0 : a;
            ob = z ? (C)0 : b;
        } else {
            oa = x0 > nx - 2 ? b : a;  // right of the grid: column nx-1 twice
            ob = x0 < 0 ? a : b;       // left of the grid: column 0 twice
        }
    }

    template <int LV>
    __device__ __forceinline__ C cell(int rel, int dx, int h, const C (&csa)[Q], const C (&csb)[Q], int centre) const
    {
        if constexpr (BOX) return win[LV][mq(rel)][R + dx][h];
        else {
            if (rel == centre) return h ? csb[R + dx] : csa[R + dx];
            return win[LV][mq(rel)][0][h];
        }
    }

    __device__ __forceinline__ void side_faces(C& a, C& b) const
    {
        if (!xedge) return;
        if constexpr (BC == PERKS_BC_ZERO) {
            a = pout ? (C)0 : a;
            b = pout ? (C)0 : b;
        } else {
            const C vw = read_lane(a, lane_w);
            const C ve = read_lane(b, lane_e);
            a = x0 < 0 ? vw : (x0 > nx - 2 ? ve : a);
            b = x0 < 0 ? vw : (x0 > nx - 2 ? ve : b);
        }
    }

    template <int U, int LV>
    __device__ __forceinline__ void level(int i)
    {
        if constexpr (LV < L) {
            constexpr int LV1 = LV + 1;
            constexpr int REL = U - R * LV1;
            const int q = i - R * LV1;
            C va = (C)0, vb = (C)0;
            if (q >= lo[LV1] && q < hi[LV1] && q < ny) {
                C csa[Q], csb[Q];
                if constexpr (!BOX) {
                    pair_shifts<R>(win[LV][mq(REL)][0][0], win[LV][mq(REL)][0][1], csa, csb);
                }
                va = eval_shape<SHAPE, C, HAS_DIV>(k, kdiv, kd, [&](int dy, int dx) -> C {
                    return cell<LV>(REL + dy, dx, 0, csa, csb, REL);
                });
                vb = eval_shape<SHAPE, C, HAS_DIV>(k, kdiv, kd, [&](int dy, int dx) -> C {
                    return cell<LV>(REL + dy, dx, 1, csa, csb, REL);
                });
                Pair2<S>::q2(va, vb, sr, (uint32_t)q * (uint32_t)nx + (uint32_t)x0, key[LV1]);
                if constexpr (LV1 == L) {
                    if (st_ok) Pair2<S>::st(dst + (size_t)q * nx, (size_t)x0, va, vb);
                }
            }
            if constexpr (LV1 < L) {
                side_faces(va, vb);
                C row[NS][2];
                if constexpr (BOX) {
                    C sa[Q], sb[Q];
                    pair_shifts<R>(va, vb, sa, sb);
#pragma unroll
                    for (int t = 0; t < Q; ++t) {
                        row[t][0] = sa[t];
                        row[t][1] = sb[t];
                    }
                } else {
                    row[0][0] = va;
                    row[0][1] = vb;
                }
                const int jb = q - (ny - 1);
#pragma unroll
                for (int jj = 1; jj <= R; ++jj)
#pragma unroll
                    for (int t = 0; t < NS; ++t)
#pragma unroll
                        for (int h = 0; h < 2; ++h)
                            row[t][h] = jb == jj ? (BC == PERKS_BC_ZERO ? (C)0 : win[LV1][mq(REL - jj)][t][h])
                                                 : row[t][h];
#pragma unroll
                for (int t = 0; t < NS; ++t) {
                    win[LV1][mq(REL)][t][0] = row[t][0];
                    win[LV1][mq(REL)][t][1] = row[t][1];
                }
#pragma unroll
                for (int j = 1; j <= R; ++j)
#pragma unroll
                    for (int t = 0; t < NS; ++t)
#pragma unroll
                        for (int h = 0; h < 2; ++h)
                            win[LV1][mq(REL - j)][t][h] =
                                q == 0 ? (BC == PERKS_BC_ZERO ? (C)0 : row[t][h]) : win[LV1][mq(REL - j)][t][h];
            }
            level<U, LV + 1>(i);
        }
    }

    template <int U>
    __device__ __forceinline__ void stage(int i)
    {
        const Raw w = fl[U];
        fl[U] = load(i + D);  // issued before row i is waited for
        C a, b;
        arrive(i, w, a, b);
        if constexpr (BOX) {
            C sa[Q], sb[Q];
            pair_shifts<R>(a, b, sa, sb);
#pragma unroll
            for (int t = 0; t < Q; ++t) {
                win[0][mq(U)][t][0] = sa[t];
                win[0][mq(U)][t][1] = sb[t];
            }
        } else {
            win[0][mq(U)][0][0] = a;
            win[0][mq(U)][0][1] = b;
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

    // Strip whose lane 0 holds columns xs0, xs0+1 (xs0 even); stores columns [vx0, vx1).
    __device__ void run(int lane, int xs0, int vx0, int vx1)
    {
        x0 = xs0 + 2 * lane;
        pout = x0 < 0 || x0 > nx - 2;
        xedge = xs0 < 0 || xs0 + 127 > nx - 1;
        lane_w = min(max(-xs0 / 2, 0), 63);
        lane_e = min(max((nx - 2 - xs0) / 2, 0), 63);
        st_ok = x0 >= vx0 && x0 < vx1;
#pragma unroll
        for (int l = 0; l <= L; ++l) {
            lo[l] = max(0, ys - R * (L - l));
            hi[l] = l == L ? ye : min(ny + R, ye + R * (L - l));
        }
        const int i0 = ys - R * L;
        const int iend = ye - 1 + R * L;
#pragma unroll
        for (int t = 0; t < D; ++t) fl[t] = load(i0 + t);
        for (int i = i0; trip<0>(i, iend); i += UF) {
        }
    }
};

template <int SHAPE, typename S, int BC, bool HAS_DIV, bool PAIR>
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
        // Strips use the full-pass ghost rim in every pass.
        const int VW = PAIR ? 128 - 2 * pair_ghost(R, a.levels) : 64 - 2 * R * a.levels;
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
                using St = std::conditional_t<PAIR, Strip2<SHAPE, S, BC, HAS_DIV, LL>, Strip<SHAPE, S, BC, HAS_DIV, LL>>;
                St s;
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
                s.run(lane, vx0 - (PAIR ? pair_ghost(R, LL) : R * LL), vx0, vx1);
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

template <int SHAPE, typename S, bool PAIR>
void* sptr2(int bc, bool hd)
{
    if (bc == PERKS_BC_ZERO)
        return hd ? reinterpret_cast<void*>(&j2d_stream_kernel<SHAPE, S, PERKS_BC_ZERO, true, PAIR>)
                  : reinterpret_cast<void*>(&j2d_stream_kernel<SHAPE, S, PERKS_BC_ZERO, false, PAIR>);
    return hd ? reinterpret_cast<void*>(&j2d_stream_kernel<SHAPE, S, PERKS_BC_CLAMP, true, PAIR>)
              : reinterpret_cast<void*>(&j2d_stream_kernel<SHAPE, S, PERKS_BC_CLAMP, false, PAIR>);
}

template <typename S, bool PAIR>
void* sptr_p(int shape, int bc, bool hd)
{
    switch (shape) {
    case PERKS_SHAPE_2D_STAR_R1: return sptr2<PERKS_SHAPE_2D_STAR_R1, S, PAIR>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R2: return sptr2<PERKS_SHAPE_2D_STAR_R2, S, PAIR>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R3: return sptr2<PERKS_SHAPE_2D_STAR_R3, S, PAIR>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R4: return sptr2<PERKS_SHAPE_2D_STAR_R4, S, PAIR>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R5: return sptr2<PERKS_SHAPE_2D_STAR_R5, S, PAIR>(bc, hd);
    case PERKS_SHAPE_2D_STAR_R6: return sptr2<PERKS_SHAPE_2D_STAR_R6, S, PAIR>(bc, hd);
    case PERKS_SHAPE_2D_BOX_R1: return sptr2<PERKS_SHAPE_2D_BOX_R1, S, PAIR>(bc, hd);
    case PERKS_SHAPE_2D_BOX_R2: return sptr2<PERKS_SHAPE_2D_BOX_R2, S, PAIR>(bc, hd);
    default: return nullptr;
    }
}

// pair: the two-columns-per-lane form (FP32, FP16 and BF16 storage only).
template <typename S>
void* sptr(int shape, int bc, bool hd, bool pair)
{
    if constexpr (std::is_same_v<S, double>) {
        return pair ? nullptr : sptr_p<S, false>(shape, bc, hd);
    } else {
        return pair ? sptr_p<S, true>(shape, bc, hd) : sptr_p<S, false>(shape, bc, hd);
    }
}

}  // namespace j2ds
}  // namespace perks
