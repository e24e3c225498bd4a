// j3d_persistent_kernel.h -- the persistent 3D 7-point kernel template.
//
// Instantiated per element type in j3d_persistent_f64.hip / _f32.hip (so the
// two build in parallel); the host side (planning, launch) is in
// j3d_persistent.hip.  The design is described there.
#pragma once

#include <hip/hip_runtime.h>

#include "grid_sync.h"
#include "j3d_common.h"
#include "storage16.h"

namespace perks {
namespace j3dp {

constexpr int kMaxLevels = 4;  // time steps per pass the kernel is built for

// Forms of the 7-point update (template argument UP):
//   kUpdGold    star1_update: every product and sum rounded as the gold writes them
//   kUpdFused   star1_update_fused: products after the first fused into the sum
//   kUpdShared  the gold's bits with fewer multiplies, for six equal positive
//               neighbour weights k: every neighbour term is k times a cell
//               value, so k*x is formed once per cell of a row (and of the
//               rows just outside it) and shifted or indexed to every term
//               that reads x, instead of once per term
constexpr int kUpdGold = 0;
constexpr int kUpdFused = 1;
constexpr int kUpdShared = 2;

struct PersistArgs {
    int nz, ny, nx;
    int steps;
    int levels;                 // time steps per pass (L of the full passes)
    int nzs[kMaxLevels + 1];    // z-slabs per pass with l levels
    GridSync* sync;             // grid barrier state (grid_sync.h)
    unsigned long long timeout; // barrier wait limit, wall-clock ticks
    uint32_t seed;              // stochastic rounding of 16-bit storage
};

template <typename T, int BC>
__device__ __forceinline__ T bc_value(T c)
{
    return BC == PERKS_BC_CLAMP ? c : (T)0;
}

// Storage format of the field: element type S in HBM, arithmetic type T.
// FP16/BF16 values are widened on load and every level's result is rounded
// to the storage format (the value a stored field would hold after that time
// step), so a pass of L levels equals L stored steps bit for bit.
template <int FMT>
struct Elem;
template <>
struct Elem<PERKS_F64> {
    using S = double;
    using T = double;
    __device__ static T load(rsrc_t r, uint32_t off) { return buf_load(r, off, double{}); }
    __device__ static void store(rsrc_t r, uint32_t off, T v) { buf_store(r, off, v); }
    __device__ static T q(T v) { return v; }
    template <bool SR>
    __device__ static T qr(T v, uint32_t, uint32_t) { return v; }
};
template <>
struct Elem<PERKS_F32> {
    using S = float;
    using T = float;
    __device__ static T load(rsrc_t r, uint32_t off) { return buf_load(r, off, float{}); }
    __device__ static void store(rsrc_t r, uint32_t off, T v) { buf_store(r, off, v); }
    __device__ static T q(T v) { return v; }
    template <bool SR>
    __device__ static T qr(T v, uint32_t, uint32_t) { return v; }
};
template <int FMT>
struct Elem16 {
    using S = uint16_t;
    using T = float;
    __device__ static T load(rsrc_t r, uint32_t off)
    {
        return Store16<FMT>::widen((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0));
    }
    __device__ static void store(rsrc_t r, uint32_t off, T v)
    {
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)Store16<FMT>::narrow(v), r, (int)off, 0, 0);
    }
    __device__ static T q(T v) { return round16<FMT>(v); }
    // The stored value of cell idx at the step with hash key `key`.
    template <bool SR>
    __device__ static T qr(T v, uint32_t idx, uint32_t key)
    {
        return Store16<FMT>::widen(narrow_as<FMT, SR>(v, idx, key));
    }
};
template <>
struct Elem<PERKS_F16> : Elem16<PERKS_F16> {};
template <>
struct Elem<PERKS_BF16> : Elem16<PERKS_BF16> {};

// Tiling of the x axis for L levels (one column per lane).  A tile's 64
// lanes store 64 - 2L columns: each level's valid region shrinks by a lane
// per side.  Under the zero boundary the cells beyond an x face are 0 at
// every level, which is exactly what the DPP shifts bring into lanes 0 and
// 63, so a tile whose first (last) lane is the face column loses no lane on
// that side: the two edge tiles start at the faces and store 64 - L columns
// each.  At 512 columns and L = 4 that is 60 + 7 x 56 + 60, nine tiles
// instead of ten.
__host__ __device__ inline bool x_face_tiles(int bc, int nx, int L)
{
    return bc == PERKS_BC_ZERO && nx >= 2 * (64 - L);
}
__host__ __device__ inline int x_tiles(int bc, int nx, int L)
{
    const int VW = 64 - 2 * L;
    if (!x_face_tiles(bc, nx, L)) return (nx + VW - 1) / VW;
    const int EW = 64 - L;
    return 2 + (nx - 2 * EW + VW - 1) / VW;
}
// Tile tx of ntx: stored columns [vx0, vx1) and the column of lane 0.
__host__ __device__ inline void x_tile(int bc, int nx, int L, int tx, int ntx, int& vx0, int& vx1, int& xw0)
{
    const int VW = 64 - 2 * L;
    if (!x_face_tiles(bc, nx, L)) {
        vx0 = tx * VW;
        vx1 = vx0 + VW < nx ? vx0 + VW : nx;
        xw0 = vx0 - L;
        return;
    }
    const int EW = 64 - L;
    if (tx == 0) {
        vx0 = 0;
        vx1 = EW;
        xw0 = 0;
    } else if (tx == ntx - 1) {
        vx0 = nx - EW;
        vx1 = nx;
        xw0 = nx - 64;
    } else {
        vx0 = EW + (tx - 1) * VW;
        vx1 = vx0 + VW < nx - EW ? vx0 + VW : nx - EW;
        xw0 = vx0 - L;
    }
}

// Tiling of the y axis for L levels: a workgroup's `rows` rows store
// VH = rows - 2(L-1), the outermost L-1 rows on each side being ghosts at
// level L.  Rows on a y face apply the boundary rule by selects at every
// level, whatever lies beyond, so a tile whose first (last) row is the face
// row loses no row on that side: the two edge tiles start at the faces and
// store rows - (L-1) rows each.  At 512 rows, 40 rows per tile and L = 4 that
// is 37 + 13 x 34 + 37, fifteen tiles instead of sixteen.
__host__ __device__ inline bool y_face_tiles(int ny, int rows, int L)
{
    return ny >= 2 * (rows - (L - 1));
}
__host__ __device__ inline int y_tiles(int ny, int rows, int L)
{
    const int VH = rows - 2 * (L - 1);
    if (!y_face_tiles(ny, rows, L)) return (ny + VH - 1) / VH;
    const int EH = rows - (L - 1);
    return 2 + (ny - 2 * EH + VH - 1) / VH;
}
// Tile ty of nty: stored rows [vy0, vy1) and the row of the workgroup's first row.
__host__ __device__ inline void y_tile(int ny, int rows, int L, int ty, int nty, int& vy0, int& vy1, int& yw0)
{
    const int VH = rows - 2 * (L - 1);
    if (!y_face_tiles(ny, rows, L)) {
        vy0 = ty * VH;
        vy1 = vy0 + VH < ny ? vy0 + VH : ny;
        yw0 = vy0 - (L - 1);
        return;
    }
    const int EH = rows - (L - 1);
    if (ty == 0) {
        vy0 = 0;
        vy1 = EH;
        yw0 = 0;
    } else if (ty == nty - 1) {
        vy0 = ny - EH;
        vy1 = ny;
        yw0 = ny - rows;
    } else {
        vy0 = EH + (ty - 1) * VH;
        vy1 = vy0 + VH < ny - EH ? vy0 + VH : ny - EH;
        yw0 = vy0 - (L - 1);
    }
}

// Work items of one pass: item = t + ntiles * (z-slab) marches planes
// [zs, ze) of x-y tile t, one of nzs equal z-slabs.  Items are dealt to the
// workgroups in turn, so at any moment the grid marches neighbouring tiles
// over neighbouring planes and the ghost rims one tile reads from HBM are
// still in L2 / the MALL when its neighbours read them.  (Cutting the
// tile-major (tile, plane) sequence into equal ranges per workgroup instead
// balances the work exactly but scatters the workgroups over z: 33% slower
// at 512^3 FP64.)
__device__ __forceinline__ void item_piece(int item, int ntiles, int nzs, int nz, int& t, int& zs, int& ze)
{
    t = item % ntiles;
    const int zsl = item / ntiles;
    zs = (int)(((long long)zsl * nz) / nzs);
    ze = (int)(((long long)(zsl + 1) * nz) / nzs);
}

// One work item: L time steps over an x-y tile and a z-slab.
//
// Lanes 0 and 63 are ghost lanes at level 0 already (no halo column loads):
// a tile stores VW = 64 - 2L columns.  Rows: each wave loads its NY rows plus
// one halo row on each side at level 0; from level 1 on, the rows that cross
// between waves come from the LDS exchange, and the workgroup's outermost
// rows are ghosts, so a tile stores VH = NW*NY - 2(L-1) rows.
//
// Level l at iteration i computes plane i-l from level-(l-1) planes i-l-1,
// i-l and i-l+1.  Of these only the centre plane i-l needs the rows outside
// the wave's strip, and it was computed one iteration earlier: every level
// publishes its two edge rows to LDS (buffer parity i&1) and the level above
// reads the neighbours' rows of its centre plane from the other parity one
// iteration later.  So the rings keep NY rows per plane, and one workgroup
// barrier per iteration, after all levels, separates the writes of one
// parity from its reads and the reads from the next overwrite.
template <typename T>
__device__ __forceinline__ T read_lane_t(T v, int l);
template <>
__device__ __forceinline__ float read_lane_t<float>(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <>
__device__ __forceinline__ double read_lane_t<double>(double v, int l)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Side faces: the lanes whose columns lie beyond an x face of the domain hold,
// at every level, what the boundary rule reads there (0, or the edge column:
// level 0 through the load, higher levels through side_faces), so waves at
// an x face take the update without boundary tests.  Waves at a y face and
// the planes at z faces apply the rule by selects (compute<true>).
template <int FMT, int BC, int NY, int NW, int P, int L, bool HAS_DIV, bool SR, int UP>
struct March {
    using E = Elem<FMT>;
    using S = typename E::S;
    using T = typename E::T;
    static constexpr int R = NY + 2;  // level-0 rows held per plane: the strip and one halo row each side
    static constexpr int D = P - 3;   // level-0 planes loaded ahead of the one level 1 needs next
    static constexpr int UF = P % 3 == 0 ? P : 3 * P;  // stages per loop trip (ring slots are static)
    const S* __restrict__ src;
    S* __restrict__ dst;
    T* xch;                           // LDS exchange rows [2][kMaxLevels-1][NW][2][64]
    int nz, ny, nx;
    size_t plane;
    uint32_t plane_bytes;
    int lane, wave;
    T k[8];
    // per item
    uint32_t roff[R];
    uint32_t soff[NY];
    bool xlo, xhi;
    bool ylo[NY], yhi[NY];
    bool edge;          // this wave's strip touches a y face of the domain
    int xcol;           // this lane's column
    bool xedge;         // some lanes lie beyond a side face
    int lane_w, lane_e; // lanes of columns 0 and nx-1 (clamped into the wave)
    int lo[L + 1], hi[L + 1];
    int base;
    // stochastic rounding: index of (y0, x) within a plane, cells per plane,
    // and the hash key of each level's time step
    uint32_t cbase, plane32;
    uint32_t key[L + 1];

    T w0[P][R];
    T wl[L > 1 ? L - 1 : 1][3][NY];

    __device__ __forceinline__ void load0(int q, int slot)
    {
        const int qc = q < nz ? q : nz - 1;
        const rsrc_t rs = make_rsrc(src + (size_t)qc * plane, plane_bytes);
#pragma unroll
        for (int r = 0; r < R; ++r) w0[slot][r] = E::load(rs, roff[r]);
    }

    // Level-lv plane p from the level-(lv-1) planes b (p-1), c (p) and u
    // (p+1), given by the wave's NY rows each (at offset `o` in their arrays)
    // and the two rows of c just outside the strip.  EDGE adds the boundary
    // rule at domain faces; interior waves skip it.
    template <bool EDGE, int RB>
    __device__ __forceinline__ void compute(const T (&b)[RB], const T (&c)[RB], const T (&u)[RB], int o, T c_lo,
                                            T c_hi, int p, int lv, T (&out)[NY]) const
    {
        const bool zlo = EDGE && p == 0;
        const bool zhi = EDGE && p == nz - 1;
        if constexpr (UP == kUpdShared) {
            // q[j] = k * (row j-1 of the centre plane): the north, south, east
            // and west terms of the neighbouring rows and lanes.  A lane with
            // no neighbour gets +0 from the shift, which is k * 0 for k > 0,
            // as in the gold form.
            const T kn = k[0];
            T q[NY + 2];
            q[0] = kn * c_lo;
#pragma unroll
            for (int r = 0; r < NY; ++r) q[r + 1] = kn * c[o + r];
            q[NY + 1] = kn * c_hi;
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                const T cc = c[o + r];
                T te = dpp_shift0<kDppWaveShl1>(q[r + 1]);
                T tw = dpp_shift0<kDppWaveShr1>(q[r + 1]);
                T tn = q[r + 2], ts = q[r];
                T tt = kn * u[o + r], tb = kn * b[o + r];
                if (EDGE) {
                    const T bq = BC == PERKS_BC_CLAMP ? q[r + 1] : (T)0;  // k * (the boundary value)
                    tw = xlo ? bq : tw;
                    te = xhi ? bq : te;
                    tn = yhi[r] ? bq : tn;
                    ts = ylo[r] ? bq : ts;
                    tt = zhi ? bq : tt;
                    tb = zlo ? bq : tb;
                }
                T acc = te + tw;
                acc = acc + tn;
                acc = acc + ts;
                acc = acc + tt;
                acc = acc + tb;
                acc = acc + k[6] * cc;
                if (HAS_DIV) acc = acc / k[7];
                out[r] = E::template qr<SR>(acc, (uint32_t)p * plane32 + cbase + (uint32_t)(r * nx), key[lv]);
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const T cc = c[o + r];
            T w = dpp_shift0<kDppWaveShr1>(cc);
            T e = dpp_shift0<kDppWaveShl1>(cc);
            T n = r + 1 < NY ? c[o + r + 1] : c_hi;
            T s = r > 0 ? c[o + r - 1] : c_lo;
            T t = u[o + r], bt = b[o + r];
            if (EDGE) {
                const T bv = bc_value<T, BC>(cc);
                w = xlo ? bv : w;
                e = xhi ? bv : e;
                n = yhi[r] ? bv : n;
                s = ylo[r] ? bv : s;
                t = zhi ? bv : t;
                bt = zlo ? bv : bt;
            }
            const T v = UP == kUpdFused ? star1_update_fused<T, HAS_DIV>(k, e, w, n, s, t, bt, cc)
                                        : star1_update<T, HAS_DIV>(k, e, w, n, s, t, bt, cc);
            out[r] = E::template qr<SR>(v, (uint32_t)p * plane32 + cbase + (uint32_t)(r * nx), key[lv]);
        }
    }

    template <int RB>
    __device__ __forceinline__ void level(const T (&b)[RB], const T (&c)[RB], const T (&u)[RB], int o, T c_lo,
                                          T c_hi, int p, int lv, T (&out)[NY]) const
    {
        if (edge || p == 0 || p == nz - 1) compute<true>(b, c, u, o, c_lo, c_hi, p, lv, out);
        else compute<false>(b, c, u, o, c_lo, c_hi, p, lv, out);
    }

    // Lanes beyond a side face take the boundary value of their row.
    __device__ __forceinline__ void side_faces(T (&out)[NY]) const
    {
        if (!xedge) return;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            if constexpr (BC == PERKS_BC_ZERO) {
                out[r] = (xcol < 0 || xcol > nx - 1) ? (T)0 : out[r];
            } else {
                const T vw = read_lane_t<T>(out[r], lane_w);
                const T ve = read_lane_t<T>(out[r], lane_e);
                out[r] = xcol < 0 ? vw : (xcol > nx - 1 ? ve : out[r]);
            }
        }
    }

    __device__ __forceinline__ T* xrows(int par, int lv) const
    {
        return xch + ((size_t)(par * (kMaxLevels - 1) + lv - 1) * NW) * 2 * 64;
    }

    // Publishes rows 0 and NY-1 of a level-lv plane (buffer parity par) and
    // keeps the plane's NY rows in the ring.
    __device__ __forceinline__ void publish(int lv, int par, const T (&out)[NY], T (&ring)[NY]) const
    {
        T* x = xrows(par, lv);
        x[(wave * 2 + 0) * 64 + lane] = out[0];
        x[(wave * 2 + 1) * 64 + lane] = out[NY - 1];
#pragma unroll
        for (int r = 0; r < NY; ++r) ring[r] = out[r];
    }

    // The rows just below and above this wave's strip of the level-lv plane
    // published with parity par.  The outermost waves read their own rows:
    // they are ghosts there.
    __device__ __forceinline__ void fetch(int lv, int par, T& c_lo, T& c_hi) const
    {
        const T* x = xrows(par, lv);
        const int wb = wave > 0 ? wave - 1 : 0;
        const int wa = wave < NW - 1 ? wave + 1 : NW - 1;
        c_lo = x[(wb * 2 + 1) * 64 + lane];
        c_hi = x[(wa * 2 + 0) * 64 + lane];
    }

    __device__ __forceinline__ void store(int p, const T (&out)[NY]) const
    {
        const rsrc_t rs = make_rsrc(dst + (size_t)p * plane, plane_bytes);
#pragma unroll
        for (int r = 0; r < NY; ++r) E::store(rs, soff[r], out[r]);
    }

    static constexpr int md(int a) { return ((a % 3) + 3) % 3; }
    static constexpr int mp(int a) { return ((a % P) + P) % P; }

    // Levels 2..L of iteration i (stage U = (i - base) mod UF); hl holds the
    // halo rows of each level's centre plane, read at the start of the stage.
    template <int U, int LV>
    __device__ __forceinline__ void upper(int i, const T (&hl)[L > 1 ? L - 1 : 1][2])
    {
        if constexpr (LV <= L) {
            const int p = i - LV;
            if (p >= lo[LV] && p < hi[LV]) {
                const T c_lo = hl[LV - 2][0], c_hi = hl[LV - 2][1];
                T out[NY];
                level(wl[LV - 2][md(U - LV - 1)], wl[LV - 2][md(U - LV)], wl[LV - 2][md(U - LV + 1)], 0, c_lo, c_hi,
                      p, LV, out);
                if constexpr (LV == L) {
                    store(p, out);
                } else {
                    side_faces(out);
                    publish(LV, i & 1, out, wl[LV - 1][md(U - LV)]);
                }
            }
            upper<U, LV + 1>(i, hl);
        }
    }

    // Iteration i: level-0 plane i is the newest one level 1 needs; plane
    // i+D is requested now into the slot plane i-3 (no longer needed) held.
    template <int U>
    __device__ __forceinline__ void stage(int i)
    {
        load0(i + D, mp(U + D));
        // The halo rows every upper level needs were published last iteration:
        // request them all now, so they arrive while level 1 computes.
        T hl[L > 1 ? L - 1 : 1][2];
#pragma unroll
        for (int lv = 2; lv <= L; ++lv) fetch(lv - 1, (i - 1) & 1, hl[lv - 2][0], hl[lv - 2][1]);
        const int p = i - 1;
        if (p >= lo[1] && p < hi[1]) {
            T out[NY];
            const T(&c)[R] = w0[mp(U - 1)];
            level(w0[mp(U - 2)], c, w0[mp(U)], 1, c[0], c[NY + 1], p, 1, out);
            if constexpr (L == 1) {
                store(p, out);
            } else {
                side_faces(out);
                publish(1, i & 1, out, wl[0][md(U - 1)]);
            }
        }
        upper<U, 2>(i, hl);
        if constexpr (L > 1) lds_barrier();
    }
    template <int U>
    __device__ __forceinline__ bool trip(int i, int iend)
    {
        if constexpr (U < UF) {
            if (i > iend) return false;
            stage<(U + 1) % UF>(i);
            return trip<U + 1>(i + 1, iend);
        }
        return true;
    }

    // Planes [zs, ze) of tile (tx, ty) of ntx x nty.
    __device__ __forceinline__ void run(int tx, int ty, int ntx, int nty, int zs, int ze)
    {
        // The exchange rows of the previous piece may still be being read.
        lds_barrier();
        int vx0, vx1, xw0;  // stored columns, column of lane 0
        x_tile(BC, nx, L, tx, ntx, vx0, vx1, xw0);
        int vy0, vy1, yw0;  // stored rows, row of wave 0's first row
        y_tile(ny, NW * NY, L, ty, nty, vy0, vy1, yw0);
        const int x = xw0 + lane;
        const int xc = x < 0 ? 0 : (x >= nx ? nx - 1 : x);
        const int y0 = yw0 + wave * NY;
        xlo = x == 0;
        xhi = x == nx - 1;
        xcol = x;
        cbase = (uint32_t)y0 * (uint32_t)nx + (uint32_t)x;
        const bool xout = x >= vx0 && x < vx1;
        xedge = xw0 < 0 || xw0 + 63 > nx - 1;
        lane_w = min(max(-xw0, 0), 63);
        lane_e = min(max(nx - 1 - xw0, 0), 63);
        edge = y0 <= 0 || y0 + NY - 1 >= ny - 1;
        // Level 0 outside the domain: the clamped cell (clamp) or 0 (zero, an
        // out-of-range buffer offset).
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int yy = y0 + r - 1;
            const int yc = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
            const bool in = yy >= 0 && yy < ny && x >= 0 && x < nx;
            roff[r] = (BC == PERKS_BC_ZERO && !in) ? kOob : (uint32_t)((yc * nx + xc) * (int)sizeof(S));
        }
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int y = y0 + r;
            ylo[r] = y == 0;
            yhi[r] = y == ny - 1;
            soff[r] = (xout && y >= vy0 && y < vy1) ? (uint32_t)((y * nx + x) * (int)sizeof(S)) : kOob;
        }
#pragma unroll
        for (int l = 0; l <= L; ++l) {
            lo[l] = max(0, zs - (L - l));
            hi[l] = min(nz, ze + (L - l));
        }
        base = lo[0];
#pragma unroll
        for (int q = 0; q <= D; ++q) load0(base + q, q);
        const int iend = hi[L] - 1 + L;
        for (int i = base + 1; trip<0>(i, iend); i += UF) {
        }
    }
};

// ---- two columns per lane (FP32 arithmetic) ----------------------------------------
// Storage access to the cells of an even column x0 and x0 + 1 in one buffer
// operation, and their rounding to the storage format.
template <int FMT>
struct Elem2;
template <>
struct Elem2<PERKS_F32> {
    __device__ static void load(rsrc_t r, uint32_t off, float& a, float& b)
    {
        const uint64_t u = (uint64_t)__double_as_longlong(buf_load(r, off, double{}));
        a = __uint_as_float((uint32_t)u);
        b = __uint_as_float((uint32_t)(u >> 32));
    }
    __device__ static void store(rsrc_t r, uint32_t off, float a, float b)
    {
        const uint64_t u = ((uint64_t)__float_as_uint(b) << 32) | __float_as_uint(a);
        buf_store(r, off, __longlong_as_double((long long)u));
    }
    template <bool SR>
    __device__ static void qr(float&, float&, uint32_t, uint32_t) {}
};
template <int FMT>
struct Elem2x16 {
    __device__ static void load(rsrc_t r, uint32_t off, float& a, float& b)
    {
        const uint32_t w = __float_as_uint(buf_load(r, off, float{}));
        a = widen_lo<FMT>(w);
        b = widen_hi<FMT>(w);
    }
    // a and b are storage values already: the conversion is exact.
    __device__ static void store(rsrc_t r, uint32_t off, float a, float b)
    {
        buf_store(r, off, __uint_as_float(Store16<FMT>::narrow2(a, b)));
    }
    template <bool SR>
    __device__ static void qr(float& a, float& b, uint32_t idx0, uint32_t key)
    {
        if constexpr (SR) {
            a = Store16<FMT>::widen(narrow_sr<FMT>(a, lowbias32(idx0 ^ key)));
            b = Store16<FMT>::widen(narrow_sr<FMT>(b, lowbias32((idx0 + 1u) ^ key)));
        } else {
            const uint32_t w = Store16<FMT>::narrow2(a, b);
            a = widen_lo<FMT>(w);
            b = widen_hi<FMT>(w);
        }
    }
};
template <>
struct Elem2<PERKS_F16> : Elem2x16<PERKS_F16> {};
template <>
struct Elem2<PERKS_BF16> : Elem2x16<PERKS_BF16> {};

// March with each lane holding two neighbouring columns (a: even x0, b: x0+1),
// so one load, store, LDS access and pair of lane shifts serves two cells.
// A tile is 128 columns wide with a ghost rim of G = L rounded up to even
// columns (stores stay on whole pairs; the grid width must be even).  Level
// structure, the one-iteration-late row exchange and the face handling are
// those of March.
template <int FMT, int BC, int NY, int NW, int P, int L, bool HAS_DIV, bool SR, int UP>
struct March2 {
    using E = Elem2<FMT>;
    using S = typename Elem<FMT>::S;
    using T = float;
    static constexpr int R = NY + 2;
    static constexpr int D = P - 3;
    static constexpr int UF = P % 3 == 0 ? P : 3 * P;
    static constexpr int G = 2 * ((L + 1) / 2);
    const S* __restrict__ src;
    S* __restrict__ dst;
    float2* xch;  // LDS exchange rows [2][kMaxLevels-1][NW][2][64] of pairs
    int nz, ny, nx;
    size_t plane;
    uint32_t plane_bytes;
    int lane, wave;
    T k[8];
    uint32_t roff[R];
    uint32_t soff[NY];
    bool xlo, xhi;  // this pair holds column 0 (half a) / nx-1 (half b)
    bool ylo[NY], yhi[NY];
    bool edge;
    int x0;
    bool xedge, pout;
    int lane_w, lane_e;
    int lo[L + 1], hi[L + 1];
    int base;
    uint32_t cbase, plane32;
    uint32_t key[L + 1];

    T w0[P][R][2];
    T wl[L > 1 ? L - 1 : 1][3][NY][2];

    __device__ __forceinline__ void load0(int q, int slot)
    {
        const int qc = q < nz ? q : nz - 1;
        const rsrc_t rs = make_rsrc(src + (size_t)qc * plane, plane_bytes);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            T a, b;
            E::load(rs, roff[r], a, b);
            if constexpr (BC == PERKS_BC_CLAMP) {
                // A pair beyond a side face loads the edge pair: both halves take the edge column.
                w0[slot][r][0] = x0 > nx - 2 ? b : a;
                w0[slot][r][1] = x0 < 0 ? a : b;
            } else {
                w0[slot][r][0] = a;
                w0[slot][r][1] = b;
            }
        }
    }

    template <bool EDGE, int RB>
    __device__ __forceinline__ void compute(const T (&b)[RB][2], const T (&c)[RB][2], const T (&u)[RB][2], int o,
                                            const T (&c_lo)[2], const T (&c_hi)[2], int p, int lv,
                                            T (&out)[NY][2]) const
    {
        const bool zlo = EDGE && p == 0;
        const bool zhi = EDGE && p == nz - 1;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const T ca = c[o + r][0], cb = c[o + r][1];
            // column x0-1 is half b of the lane before; column x0+2 half a of the lane after
            T wa = dpp_shift0<kDppWaveShr1>(cb), ea = cb;
            T wb = ca, eb = dpp_shift0<kDppWaveShl1>(ca);
            T na = r + 1 < NY ? c[o + r + 1][0] : c_hi[0], nb = r + 1 < NY ? c[o + r + 1][1] : c_hi[1];
            T sa = r > 0 ? c[o + r - 1][0] : c_lo[0], sb = r > 0 ? c[o + r - 1][1] : c_lo[1];
            T ta = u[o + r][0], tb = u[o + r][1], ba = b[o + r][0], bb = b[o + r][1];
            if (EDGE) {
                const T bva = bc_value<T, BC>(ca), bvb = bc_value<T, BC>(cb);
                wa = xlo ? bva : wa;
                eb = xhi ? bvb : eb;
                na = yhi[r] ? bva : na;
                nb = yhi[r] ? bvb : nb;
                sa = ylo[r] ? bva : sa;
                sb = ylo[r] ? bvb : sb;
                ta = zhi ? bva : ta;
                tb = zhi ? bvb : tb;
                ba = zlo ? bva : ba;
                bb = zlo ? bvb : bb;
            }
            T va = UP == kUpdFused ? star1_update_fused<T, HAS_DIV>(k, ea, wa, na, sa, ta, ba, ca)
                                   : star1_update<T, HAS_DIV>(k, ea, wa, na, sa, ta, ba, ca);
            T vb = UP == kUpdFused ? star1_update_fused<T, HAS_DIV>(k, eb, wb, nb, sb, tb, bb, cb)
                                   : star1_update<T, HAS_DIV>(k, eb, wb, nb, sb, tb, bb, cb);
            E::template qr<SR>(va, vb, (uint32_t)p * plane32 + cbase + (uint32_t)(r * nx), key[lv]);
            out[r][0] = va;
            out[r][1] = vb;
        }
    }

    template <int RB>
    __device__ __forceinline__ void level(const T (&b)[RB][2], const T (&c)[RB][2], const T (&u)[RB][2], int o,
                                          const T (&c_lo)[2], const T (&c_hi)[2], int p, int lv,
                                          T (&out)[NY][2]) const
    {
        if (edge || p == 0 || p == nz - 1) compute<true>(b, c, u, o, c_lo, c_hi, p, lv, out);
        else compute<false>(b, c, u, o, c_lo, c_hi, p, lv, out);
    }

    __device__ __forceinline__ void side_faces(T (&out)[NY][2]) const
    {
        if (!xedge) return;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            if constexpr (BC == PERKS_BC_ZERO) {
                out[r][0] = pout ? (T)0 : out[r][0];
                out[r][1] = pout ? (T)0 : out[r][1];
            } else {
                const T vw = read_lane_t<T>(out[r][0], lane_w);
                const T ve = read_lane_t<T>(out[r][1], lane_e);
                out[r][0] = x0 < 0 ? vw : (x0 > nx - 2 ? ve : out[r][0]);
                out[r][1] = x0 < 0 ? vw : (x0 > nx - 2 ? ve : out[r][1]);
            }
        }
    }

    __device__ __forceinline__ float2* xrows(int par, int lv) const
    {
        return xch + ((size_t)(par * (kMaxLevels - 1) + lv - 1) * NW) * 2 * 64;
    }

    __device__ __forceinline__ void publish(int lv, int par, const T (&out)[NY][2], T (&ring)[NY][2]) const
    {
        float2* x = xrows(par, lv);
        x[(wave * 2 + 0) * 64 + lane] = make_float2(out[0][0], out[0][1]);
        x[(wave * 2 + 1) * 64 + lane] = make_float2(out[NY - 1][0], out[NY - 1][1]);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            ring[r][0] = out[r][0];
            ring[r][1] = out[r][1];
        }
    }

    __device__ __forceinline__ void fetch(int lv, int par, T (&c_lo)[2], T (&c_hi)[2]) const
    {
        const float2* x = xrows(par, lv);
        const int wb = wave > 0 ? wave - 1 : 0;
        const int wa = wave < NW - 1 ? wave + 1 : NW - 1;
        const float2 lo2 = x[(wb * 2 + 1) * 64 + lane];
        const float2 hi2 = x[(wa * 2 + 0) * 64 + lane];
        c_lo[0] = lo2.x;
        c_lo[1] = lo2.y;
        c_hi[0] = hi2.x;
        c_hi[1] = hi2.y;
    }

    __device__ __forceinline__ void store(int p, const T (&out)[NY][2]) const
    {
        const rsrc_t rs = make_rsrc(dst + (size_t)p * plane, plane_bytes);
#pragma unroll
        for (int r = 0; r < NY; ++r) E::store(rs, soff[r], out[r][0], out[r][1]);
    }

    static constexpr int md(int a) { return ((a % 3) + 3) % 3; }
    static constexpr int mp(int a) { return ((a % P) + P) % P; }

    template <int U, int LV>
    __device__ __forceinline__ void upper(int i, const T (&hl)[L > 1 ? L - 1 : 1][2][2])
    {
        if constexpr (LV <= L) {
            const int p = i - LV;
            if (p >= lo[LV] && p < hi[LV]) {
                T out[NY][2];
                level(wl[LV - 2][md(U - LV - 1)], wl[LV - 2][md(U - LV)], wl[LV - 2][md(U - LV + 1)], 0, hl[LV - 2][0],
                      hl[LV - 2][1], p, LV, out);
                if constexpr (LV == L) {
                    store(p, out);
                } else {
                    side_faces(out);
                    publish(LV, i & 1, out, wl[LV - 1][md(U - LV)]);
                }
            }
            upper<U, LV + 1>(i, hl);
        }
    }

    template <int U>
    __device__ __forceinline__ void stage(int i)
    {
        load0(i + D, mp(U + D));
        T hl[L > 1 ? L - 1 : 1][2][2];
#pragma unroll
        for (int lv = 2; lv <= L; ++lv) fetch(lv - 1, (i - 1) & 1, hl[lv - 2][0], hl[lv - 2][1]);
        const int p = i - 1;
        if (p >= lo[1] && p < hi[1]) {
            T out[NY][2];
            const T(&c)[R][2] = w0[mp(U - 1)];
            level(w0[mp(U - 2)], c, w0[mp(U)], 1, c[0], c[NY + 1], p, 1, out);
            if constexpr (L == 1) {
                store(p, out);
            } else {
                side_faces(out);
                publish(1, i & 1, out, wl[0][md(U - 1)]);
            }
        }
        upper<U, 2>(i, hl);
        if constexpr (L > 1) lds_barrier();
    }

    template <int U>
    __device__ __forceinline__ bool trip(int i, int iend)
    {
        if constexpr (U < UF) {
            if (i > iend) return false;
            stage<(U + 1) % UF>(i);
            return trip<U + 1>(i + 1, iend);
        }
        return true;
    }

    // Planes [zs, ze) of tile (tx, ty) of nty rows of tiles.
    __device__ __forceinline__ void run(int tx, int ty, int nty, int zs, int ze)
    {
        lds_barrier();
        constexpr int VW = 128 - 2 * G;
        const int vx0 = tx * VW, vx1 = min(vx0 + VW, nx);
        int vy0, vy1, yw0;
        y_tile(ny, NW * NY, L, ty, nty, vy0, vy1, yw0);
        const int xw0 = vx0 - G;  // column of lane 0's half a (even)
        x0 = xw0 + 2 * lane;
        const int xc = x0 < 0 ? 0 : (x0 > nx - 2 ? nx - 2 : x0);
        const int y0 = yw0 + wave * NY;
        xlo = x0 == 0;
        xhi = x0 + 1 == nx - 1;
        pout = x0 < 0 || x0 > nx - 2;
        cbase = (uint32_t)y0 * (uint32_t)nx + (uint32_t)x0;
        const bool xstore = x0 >= vx0 && x0 < vx1;
        xedge = xw0 < 0 || xw0 + 127 > nx - 1;
        lane_w = min(max(-xw0 / 2, 0), 63);
        lane_e = min(max((nx - 2 - xw0) / 2, 0), 63);
        edge = y0 <= 0 || y0 + NY - 1 >= ny - 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int yy = y0 + r - 1;
            const int yc = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
            const bool in = yy >= 0 && yy < ny && !pout;
            roff[r] = (BC == PERKS_BC_ZERO && !in) ? kOob : (uint32_t)((yc * nx + xc) * (int)sizeof(S));
        }
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int y = y0 + r;
            ylo[r] = y == 0;
            yhi[r] = y == ny - 1;
            soff[r] = (xstore && y >= vy0 && y < vy1) ? (uint32_t)((y * nx + x0) * (int)sizeof(S)) : kOob;
        }
#pragma unroll
        for (int l = 0; l <= L; ++l) {
            lo[l] = max(0, zs - (L - l));
            hi[l] = min(nz, ze + (L - l));
        }
        base = lo[0];
#pragma unroll
        for (int q = 0; q <= D; ++q) load0(base + q, q);
        const int iend = hi[L] - 1 + L;
        for (int i = base + 1; trip<0>(i, iend); i += UF) {
        }
    }
};

template <int FMT, int BC, int NY, int NW, int P, int L, bool HAS_DIV, bool SR, int UP>
__device__ __forceinline__ void run_pass2(const typename Elem<FMT>::S* src, typename Elem<FMT>::S* dst, float2* xch,
                                          const PersistArgs& a, const float* k, int lane, int wave, int step0)
{
    constexpr int VW = 128 - 2 * (2 * ((L + 1) / 2));
    const int ntx = (a.nx + VW - 1) / VW;
    const int nty = y_tiles(a.ny, NW * NY, L);
    const int nitems = ntx * nty * a.nzs[L];
    const int G = gridDim.x;
    for (int item = xcd_tile_id(blockIdx.x, G); item < nitems; item += G) {
        int t, zs, ze;
        item_piece(item, ntx * nty, a.nzs[L], a.nz, t, zs, ze);
        March2<FMT, BC, NY, NW, P, L, HAS_DIV, SR, UP> m;
        m.src = src;
        m.dst = dst;
        m.xch = xch;
        m.nz = a.nz;
        m.ny = a.ny;
        m.nx = a.nx;
        m.plane = (size_t)a.ny * a.nx;
        m.plane_bytes = (uint32_t)(m.plane * sizeof(typename Elem<FMT>::S));
        m.lane = lane;
        m.wave = wave;
        m.plane32 = (uint32_t)m.plane;
        if constexpr (SR) {
#pragma unroll
            for (int l = 1; l <= L; ++l) m.key[l] = sr_key((uint32_t)(step0 + l - 1), a.seed);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) m.k[i] = k[i];
        m.run(t % ntx, t / ntx, nty, zs, ze);
    }
}

template <int FMT, int BC, int NY, int NW, int P, int WPE, bool HAS_DIV, bool SR, int UP>
__global__ void __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(WPE)))
j3d_star1_persistent2_kernel(typename Elem<FMT>::S* buf_a, typename Elem<FMT>::S* buf_b, PersistArgs a, KCoefs cf)
{
    using S = typename Elem<FMT>::S;
    __shared__ float2 xch[2 * (kMaxLevels - 1) * NW * 2 * 64];
    float k[8];
    load_coefs7<float>(cf, k);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int L = a.levels;
    const int npass = (a.steps + L - 1) / L;
    S* src = buf_a;
    S* dst = buf_b;
    for (int pass = 0; pass < npass; ++pass) {
        const int lv = (pass == npass - 1) ? a.steps - pass * L : L;
        switch (lv) {
        case 1: run_pass2<FMT, BC, NY, NW, P, 1, HAS_DIV, SR, UP>(src, dst, xch, a, k, lane, wave, pass * L); break;
        case 2: run_pass2<FMT, BC, NY, NW, P, 2, HAS_DIV, SR, UP>(src, dst, xch, a, k, lane, wave, pass * L); break;
        case 3: run_pass2<FMT, BC, NY, NW, P, 3, HAS_DIV, SR, UP>(src, dst, xch, a, k, lane, wave, pass * L); break;
        default: run_pass2<FMT, BC, NY, NW, P, 4, HAS_DIV, SR, UP>(src, dst, xch, a, k, lane, wave, pass * L); break;
        }
        S* t = src;
        src = dst;
        dst = t;
        if (pass + 1 < npass && !grid_sync(a.sync, gridDim.x, a.timeout)) return;
    }
}

template <int FMT, int BC, int NY, int NW, int P, int L, bool HAS_DIV, bool SR, int UP>
__device__ __forceinline__ void run_pass(const typename Elem<FMT>::S* src, typename Elem<FMT>::S* dst,
                                         typename Elem<FMT>::T* xch, const PersistArgs& a,
                                         const typename Elem<FMT>::T* k, int lane, int wave, int step0)
{
    const int ntx = x_tiles(BC, a.nx, L);
    const int nty = y_tiles(a.ny, NW * NY, L);
    const int nitems = ntx * nty * a.nzs[L];
    const int G = gridDim.x;
    for (int item = xcd_tile_id(blockIdx.x, G); item < nitems; item += G) {
        int t, zs, ze;
        item_piece(item, ntx * nty, a.nzs[L], a.nz, t, zs, ze);
        March<FMT, BC, NY, NW, P, L, HAS_DIV, SR, UP> m;
        m.src = src;
        m.dst = dst;
        m.xch = xch;
        m.nz = a.nz;
        m.ny = a.ny;
        m.nx = a.nx;
        m.plane = (size_t)a.ny * a.nx;
        m.plane_bytes = (uint32_t)(m.plane * sizeof(typename Elem<FMT>::S));
        m.lane = lane;
        m.wave = wave;
        m.plane32 = (uint32_t)m.plane;
        if constexpr (SR) {
#pragma unroll
            for (int l = 1; l <= L; ++l) m.key[l] = sr_key((uint32_t)(step0 + l - 1), a.seed);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) m.k[i] = k[i];
        m.run(t % ntx, t / ntx, ntx, nty, zs, ze);
    }
}

template <int FMT, int BC, int NY, int NW, int P, int WPE, bool HAS_DIV, bool SR, int UP>
__global__ void __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(WPE)))
j3d_star1_persistent_kernel(typename Elem<FMT>::S* buf_a, typename Elem<FMT>::S* buf_b, PersistArgs a, KCoefs cf)
{
    using S = typename Elem<FMT>::S;
    using T = typename Elem<FMT>::T;
    __shared__ T xch[2 * (kMaxLevels - 1) * NW * 2 * 64];
    T k[8];
    load_coefs7<T>(cf, k);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int L = a.levels;
    const int npass = (a.steps + L - 1) / L;
    S* src = buf_a;
    S* dst = buf_b;
    for (int pass = 0; pass < npass; ++pass) {
        const int lv = (pass == npass - 1) ? a.steps - pass * L : L;
        switch (lv) {
        case 1: run_pass<FMT, BC, NY, NW, P, 1, HAS_DIV, SR, UP>(src, dst, xch, a, k, lane, wave, pass * L); break;
        case 2: run_pass<FMT, BC, NY, NW, P, 2, HAS_DIV, SR, UP>(src, dst, xch, a, k, lane, wave, pass * L); break;
        case 3: run_pass<FMT, BC, NY, NW, P, 3, HAS_DIV, SR, UP>(src, dst, xch, a, k, lane, wave, pass * L); break;
        default: run_pass<FMT, BC, NY, NW, P, 4, HAS_DIV, SR, UP>(src, dst, xch, a, k, lane, wave, pass * L); break;
        }
        S* t = src;
        src = dst;
        dst = t;
        if (pass + 1 < npass && !grid_sync(a.sync, gridDim.x, a.timeout)) return;
    }
}

// Tile shapes the persistent kernel is built for: rows per wave, waves per
// workgroup, level-0 ring slots, and the minimum waves per SIMD the register
// allocation must allow (1: no limit).  The level rings dominate the VGPR
// count, so the tables differ for FP64 and for FP32 arithmetic.  Variant 0 is
// the default (the fastest measured at 512^3):
//   FP32 arithmetic (FP32, FP16, BF16 storage)    FP64
//   0  16 waves x 4 rows (64 x 64 tile)            12 waves x 3 rows, 5 slots (64 x 36)
//   1  8 waves x 4 rows, 9 slots                  8 waves x 4 rows, 9 slots
//   2  8 waves x 4 rows (two workgroups per CU)   4 waves x 4 rows (two per CU)
//   3  16 waves x 2 rows                          8 waves x 5 rows (64 x 40 tile)
//   4  8 waves x 4 rows, two columns per lane     8 waves x 4 rows (64 x 32 tile)
//      (FP32: a 128 x 32 tile; even grid widths; at 512^3 25% slower than 0
//      for FP32: twice the registers halve the occupancy.  FP64: the 12-wave
//      workgroup needs three waves per SIMD, so its allocation stays within
//      168 VGPRs (five level-0 slots; six would spill): 2.7% faster than
//      8 x 5 rows at 512^3, whose ~230 VGPRs allow two waves per SIMD; 8 x 5
//      is 3% faster than 8 x 4; 6 rows would need more than 256 VGPRs)
// Selected through opts->tune[4].
struct Variant {
    int ny, nw, p, wpe;
    int pair;  // two columns per lane (March2)
};
constexpr Variant kVarF32[] = {{4, 16, 6, 1, 0}, {4, 8, 9, 1, 0}, {4, 8, 6, 1, 0}, {2, 16, 6, 1, 0}, {4, 8, 6, 1, 1}};
constexpr Variant kVarF64[] = {{3, 12, 5, 1, 0}, {4, 8, 9, 1, 0}, {4, 4, 6, 1, 0}, {5, 8, 6, 1, 0}, {4, 8, 6, 1, 0}};
constexpr int kNumVariants = 5;
constexpr int kPairVariant = 4;
__host__ __device__ constexpr Variant variant_shape(int fmt, int v)
{
    return fmt == PERKS_F64 ? kVarF64[v] : kVarF32[v];
}

template <int FMT, int BC, bool HAS_DIV, int V, bool SR, int UP>
void* kernel_ptr()
{
    constexpr Variant v = variant_shape(FMT, V);
    // the two-column form runs the shared-product weights in the gold form (the same bits)
    constexpr int UP2 = UP == kUpdShared ? kUpdGold : UP;
    if constexpr (v.pair != 0 && FMT != PERKS_F64)
        return reinterpret_cast<void*>(&j3d_star1_persistent2_kernel<FMT, BC, v.ny, v.nw, v.p, v.wpe, HAS_DIV, SR, UP2>);
    else
        return reinterpret_cast<void*>(&j3d_star1_persistent_kernel<FMT, BC, v.ny, v.nw, v.p, v.wpe, HAS_DIV, SR, UP>);
}

template <int FMT, int BC, bool HAS_DIV, bool SR, int UP>
void* pick_variant(int var)
{
    switch (var) {
    case 1: return kernel_ptr<FMT, BC, HAS_DIV, 1, SR, UP>();
    case 2: return kernel_ptr<FMT, BC, HAS_DIV, 2, SR, UP>();
    case 3: return kernel_ptr<FMT, BC, HAS_DIV, 3, SR, UP>();
    case 4: return kernel_ptr<FMT, BC, HAS_DIV, 4, SR, UP>();
    default: return kernel_ptr<FMT, BC, HAS_DIV, 0, SR, UP>();
    }
}

template <int FMT, bool SR, int UP>
void* pick_bc(int bc, bool has_div, int var)
{
    if (bc == PERKS_BC_ZERO)
        return has_div ? pick_variant<FMT, PERKS_BC_ZERO, true, SR, UP>(var)
                       : pick_variant<FMT, PERKS_BC_ZERO, false, SR, UP>(var);
    return has_div ? pick_variant<FMT, PERKS_BC_CLAMP, true, SR, UP>(var)
                   : pick_variant<FMT, PERKS_BC_CLAMP, false, SR, UP>(var);
}

// sr: stochastic rounding (16-bit storage only; FP64/FP32 ignore it).
// up: the update form (kUpd*); the fused form is built for FP64/FP32 storage
// without a divisor (nullptr otherwise), the gold and shared-product forms for
// every storage format.
template <int FMT>
void* pick_kernel(int bc, bool has_div, int var, bool sr, int up)
{
    if constexpr (FMT == PERKS_F64 || FMT == PERKS_F32) {
        if (up == kUpdFused) {
            if (has_div) return nullptr;
            return bc == PERKS_BC_ZERO ? pick_variant<FMT, PERKS_BC_ZERO, false, false, kUpdFused>(var)
                                       : pick_variant<FMT, PERKS_BC_CLAMP, false, false, kUpdFused>(var);
        }
        if (up == kUpdShared) return pick_bc<FMT, false, kUpdShared>(bc, has_div, var);
    } else {
        if (up == kUpdFused) return nullptr;
        if (up == kUpdShared)
            return sr ? pick_bc<FMT, true, kUpdShared>(bc, has_div, var) : pick_bc<FMT, false, kUpdShared>(bc, has_div, var);
        if (sr) return pick_bc<FMT, true, kUpdGold>(bc, has_div, var);
    }
    return pick_bc<FMT, false, kUpdGold>(bc, has_div, var);
}

}  // namespace j3dp

// Kernel entry points per storage format (j3d_persistent_<fmt>.hip).
void* j3d_persistent_kernel_f64(int bc, bool has_div, int variant, bool sr, int up);
void* j3d_persistent_kernel_f32(int bc, bool has_div, int variant, bool sr, int up);
void* j3d_persistent_kernel_f16(int bc, bool has_div, int variant, bool sr, int up);
void* j3d_persistent_kernel_bf16(int bc, bool has_div, int variant, bool sr, int up);

}  // namespace perks
