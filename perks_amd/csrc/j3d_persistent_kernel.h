// j3d_persistent_kernel.h -- the persistent 3D 7-point kernel template.
//
// Instantiated per element type in j3d_persistent_f64.hip / _f32.hip (so the
// two build in parallel); the host side (planning, launch) is in
// j3d_persistent.hip.  The design is described there.
#pragma once

#include <hip/hip_runtime.h>

#include "grid_sync.h"
#include "j3d_common.h"

namespace perks {
namespace j3dp {

constexpr int kMaxLevels = 4;  // time steps per pass the kernel is built for

struct PersistArgs {
    int nz, ny, nx;
    int steps;
    int levels;                 // time steps per pass (L of the full passes)
    int nzs[kMaxLevels + 1];    // z-slabs per pass with l levels
    GridSync* sync;             // grid barrier state (grid_sync.h)
    unsigned long long timeout; // barrier wait limit, wall-clock ticks
};

template <typename T, int BC>
__device__ __forceinline__ T bc_value(T c)
{
    return BC == PERKS_BC_CLAMP ? c : (T)0;
}

// One work item: L time steps over an x-y tile and a z-slab.
//
// Lanes 0 and 63 are ghost lanes at level 0 already (no halo column loads):
// a tile stores VW = 64 - 2L columns.  Rows: each wave loads its NY rows plus
// one halo row on each side at level 0; from level 1 on, the rows that cross
// between waves come from the LDS exchange, and the workgroup's outermost
// rows are ghosts, so a tile stores VH = NW*NY - 2(L-1) rows.
template <typename T, int BC, int NY, int NW, int P, int L, bool HAS_DIV>
struct March {
    static constexpr int R = NY + 2;  // rows held per plane: the strip and one halo row each side
    static constexpr int D = P - 3;   // level-0 planes loaded ahead of the one level 1 needs next
    static constexpr int UF = P % 3 == 0 ? P : 3 * P;  // stages per loop trip (ring slots are static)
    const T* __restrict__ src;
    T* __restrict__ dst;
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
    bool edge;  // this wave's strip touches an x or y face of the domain
    int lo[L + 1], hi[L + 1];
    int base;

    T w0[P][R];
    T wl[L > 1 ? L - 1 : 1][3][R];

    __device__ __forceinline__ void load0(int q, int slot)
    {
        const int qc = q < nz ? q : nz - 1;
        const rsrc_t rs = make_rsrc(src + (size_t)qc * plane, plane_bytes);
#pragma unroll
        for (int r = 0; r < R; ++r) w0[slot][r] = buf_load(rs, roff[r], T{});
    }

    // Level-l plane p from the three level-(l-1) planes b (p-1), c (p), u (p+1).
    // EDGE adds the boundary rule at domain faces; interior waves skip it.
    template <bool EDGE>
    __device__ __forceinline__ void compute(const T (&b)[R], const T (&c)[R], const T (&u)[R], int p,
                                            T (&out)[NY]) const
    {
        const bool zlo = EDGE && p == 0;
        const bool zhi = EDGE && p == nz - 1;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const T cc = c[r + 1];
            T w = dpp_shift0<kDppWaveShr1>(cc);
            T e = dpp_shift0<kDppWaveShl1>(cc);
            T n = c[r + 2], s = c[r], t = u[r + 1], bt = b[r + 1];
            if (EDGE) {
                const T bv = bc_value<T, BC>(cc);
                w = xlo ? bv : w;
                e = xhi ? bv : e;
                n = yhi[r] ? bv : n;
                s = ylo[r] ? bv : s;
                t = zhi ? bv : t;
                bt = zlo ? bv : bt;
            }
            out[r] = star1_update<T, HAS_DIV>(k, e, w, n, s, t, bt, cc);
        }
    }

    __device__ __forceinline__ void level(const T (&b)[R], const T (&c)[R], const T (&u)[R], int p,
                                          T (&out)[NY]) const
    {
        if (edge || p == 0 || p == nz - 1) compute<true>(b, c, u, p, out);
        else compute<false>(b, c, u, p, out);
    }

    // Publishes rows 0 and NY-1 of a level plane and gathers the rows just
    // outside this wave's strip from the neighbouring waves.
    __device__ __forceinline__ void exchange(int lv, int par, const T (&out)[NY], T (&dstp)[R])
    {
        T* x = xch + ((size_t)(par * (kMaxLevels - 1) + lv - 1) * NW) * 2 * 64;
        x[(wave * 2 + 0) * 64 + lane] = out[0];
        x[(wave * 2 + 1) * 64 + lane] = out[NY - 1];
        lds_barrier();
        const int wb = wave > 0 ? wave - 1 : 0;
        const int wa = wave < NW - 1 ? wave + 1 : NW - 1;
        dstp[0] = x[(wb * 2 + 1) * 64 + lane];
        dstp[NY + 1] = x[(wa * 2 + 0) * 64 + lane];
#pragma unroll
        for (int r = 0; r < NY; ++r) dstp[r + 1] = out[r];
    }

    __device__ __forceinline__ void store(int p, const T (&out)[NY]) const
    {
        const rsrc_t rs = make_rsrc(dst + (size_t)p * plane, plane_bytes);
#pragma unroll
        for (int r = 0; r < NY; ++r) buf_store(rs, soff[r], out[r]);
    }

    static constexpr int md(int a) { return ((a % 3) + 3) % 3; }
    static constexpr int mp(int a) { return ((a % P) + P) % P; }

    // Levels 2..L of iteration i (stage U = (i - base) mod UF).
    template <int U, int LV>
    __device__ __forceinline__ void upper(int i)
    {
        if constexpr (LV <= L) {
            const int p = i - LV;
            if (p >= lo[LV] && p < hi[LV]) {
                T out[NY];
                level(wl[LV - 2][md(U - LV - 1)], wl[LV - 2][md(U - LV)], wl[LV - 2][md(U - LV + 1)], p, out);
                if constexpr (LV == L) store(p, out);
                else exchange(LV, i & 1, out, wl[LV - 1][md(U - LV)]);
            }
            upper<U, LV + 1>(i);
        }
    }

    // Iteration i: level-0 plane i is the newest one level 1 needs; plane
    // i+D is requested now into the slot plane i-3 (no longer needed) held.
    template <int U>
    __device__ __forceinline__ void stage(int i)
    {
        load0(i + D, mp(U + D));
        const int p = i - 1;
        if (p >= lo[1] && p < hi[1]) {
            T out[NY];
            level(w0[mp(U - 2)], w0[mp(U - 1)], w0[mp(U)], p, out);
            if constexpr (L == 1) store(p, out);
            else exchange(1, i & 1, out, wl[0][md(U - 1)]);
        }
        upper<U, 2>(i);
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

    __device__ void run(int item, int nzs)
    {
        // The exchange rows of the previous item may still be being read.
        lds_barrier();
        constexpr int VW = 64 - 2 * L;
        const int VH = NW * NY - 2 * (L - 1);
        const int ntx = (nx + VW - 1) / VW;
        const int nty = (ny + VH - 1) / VH;
        const int tx = item % ntx;
        const int ty = (item / ntx) % nty;
        const int zsl = item / (ntx * nty);
        const int zs = (int)(((long long)zsl * nz) / nzs);
        const int ze = (int)(((long long)(zsl + 1) * nz) / nzs);

        const int vx0 = tx * VW, vx1 = min(vx0 + VW, nx);
        const int vy0 = ty * VH, vy1 = min(vy0 + VH, ny);
        const int xw0 = vx0 - L;  // column of lane 0
        const int x = xw0 + lane;
        const int xc = x < 0 ? 0 : (x >= nx ? nx - 1 : x);
        const int y0 = vy0 - (L - 1) + wave * NY;
        xlo = x == 0;
        xhi = x == nx - 1;
        const bool xout = x >= vx0 && x < vx1;
        const bool xedge = xw0 <= 0 || xw0 + 63 >= nx - 1;
        const bool yedge = y0 <= 0 || y0 + NY - 1 >= ny - 1;
        edge = xedge || yedge;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int yy = y0 + r - 1;
            const int yc = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
            roff[r] = (uint32_t)((yc * nx + xc) * (int)sizeof(T));
        }
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int y = y0 + r;
            ylo[r] = y == 0;
            yhi[r] = y == ny - 1;
            soff[r] = (xout && y >= vy0 && y < vy1) ? (uint32_t)((y * nx + x) * (int)sizeof(T)) : kOob;
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

template <typename T, int BC, int NY, int NW, int P, int L, bool HAS_DIV>
__device__ __forceinline__ void run_pass(const T* src, T* dst, T* xch, const PersistArgs& a, const T* k, int lane,
                                         int wave)
{
    constexpr int VW = 64 - 2 * L;
    const int VH = NW * NY - 2 * (L - 1);
    const int nitems = ((a.nx + VW - 1) / VW) * ((a.ny + VH - 1) / VH) * a.nzs[L];
    const int G = gridDim.x;
    const int me = xcd_tile_id(blockIdx.x, G);
    for (int item = me; item < nitems; item += G) {
        March<T, BC, NY, NW, P, L, HAS_DIV> m;
        m.src = src;
        m.dst = dst;
        m.xch = xch;
        m.nz = a.nz;
        m.ny = a.ny;
        m.nx = a.nx;
        m.plane = (size_t)a.ny * a.nx;
        m.plane_bytes = (uint32_t)(m.plane * sizeof(T));
        m.lane = lane;
        m.wave = wave;
#pragma unroll
        for (int i = 0; i < 8; ++i) m.k[i] = k[i];
        m.run(item, a.nzs[L]);
    }
}

template <typename T, int BC, int NY, int NW, int P, bool HAS_DIV>
__global__ void __launch_bounds__(NW * 64)
j3d_star1_persistent_kernel(T* buf_a, T* buf_b, PersistArgs a, KCoefs cf)
{
    __shared__ T xch[2 * (kMaxLevels - 1) * NW * 2 * 64];
    T k[8];
    load_coefs7<T>(cf, k);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int L = a.levels;
    const int npass = (a.steps + L - 1) / L;
    T* src = buf_a;
    T* dst = buf_b;
    for (int pass = 0; pass < npass; ++pass) {
        const int lv = (pass == npass - 1) ? a.steps - pass * L : L;
        switch (lv) {
        case 1: run_pass<T, BC, NY, NW, P, 1, HAS_DIV>(src, dst, xch, a, k, lane, wave); break;
        case 2: run_pass<T, BC, NY, NW, P, 2, HAS_DIV>(src, dst, xch, a, k, lane, wave); break;
        case 3: run_pass<T, BC, NY, NW, P, 3, HAS_DIV>(src, dst, xch, a, k, lane, wave); break;
        default: run_pass<T, BC, NY, NW, P, 4, HAS_DIV>(src, dst, xch, a, k, lane, wave); break;
        }
        T* t = src;
        src = dst;
        dst = t;
        if (pass + 1 < npass && !grid_sync(a.sync, gridDim.x, a.timeout)) return;
    }
}

// Tile shapes the persistent kernel is built for: rows per wave, waves per
// workgroup, level-0 ring slots.  Variant 0 is the default; variant 1 keeps
// one more plane in flight and is selectable through opts->tune[4].
struct Variant {
    int ny, nw, p;
};
constexpr Variant kVariants[] = {{4, 8, 6}, {4, 8, 9}};
constexpr int kNumVariants = 2;

template <typename T, int BC, bool HAS_DIV, int V>
void* kernel_ptr()
{
    constexpr Variant v = kVariants[V];
    return reinterpret_cast<void*>(&j3d_star1_persistent_kernel<T, BC, v.ny, v.nw, v.p, HAS_DIV>);
}

template <typename T>
void* pick_kernel(int bc, bool has_div, int var)
{
    if (bc == PERKS_BC_ZERO) {
        if (has_div) return var == 1 ? kernel_ptr<T, PERKS_BC_ZERO, true, 1>() : kernel_ptr<T, PERKS_BC_ZERO, true, 0>();
        return var == 1 ? kernel_ptr<T, PERKS_BC_ZERO, false, 1>() : kernel_ptr<T, PERKS_BC_ZERO, false, 0>();
    }
    if (has_div) return var == 1 ? kernel_ptr<T, PERKS_BC_CLAMP, true, 1>() : kernel_ptr<T, PERKS_BC_CLAMP, true, 0>();
    return var == 1 ? kernel_ptr<T, PERKS_BC_CLAMP, false, 1>() : kernel_ptr<T, PERKS_BC_CLAMP, false, 0>();
}

}  // namespace j3dp

void* j3d_persistent_kernel_f64(int bc, bool has_div, int variant);
void* j3d_persistent_kernel_f32(int bc, bool has_div, int variant);

}  // namespace perks
