// j3d_persistent.hip -- persistent 3D Jacobi: one cooperative launch runs every time step.
//
// Replaces the reference's kernel3d_general / kernel3d_general_inner
// (stencil/3dstencil/j3d-general-wrapper.cu:23-49, j3d-general-kernels.cuh:224-1231)
// and its cooperative launch (j3d.cu:742).  The reference keeps part of each
// block's tile in registers and shared memory across grid.sync()s.  On MI355X
// a 512^3 FP64 field (1.07 GB) is six times the whole chip's LDS + VGPR
// capacity, so residency alone removes little traffic; this kernel instead
// keeps every tile resident on chip for L consecutive time steps:
//
//   * The domain is cut into work items: an x-y tile (64 lanes by NW*NY rows,
//     including a ghost rim of L-1 cells) times a z-slab.
//   * A workgroup marches through its item's slab plane by plane.  Level 0 is
//     the field as read from HBM; level l (1..L) is the field after l steps.
//     At each plane, level l computes plane p-l from three planes of level l-1
//     that it holds in VGPRs (a ring of three plane buffers per level), so a
//     plane is read from HBM once and written once per L steps.
//   * East/west neighbours come from the adjacent lane by DPP whole-wave
//     shifts, north/south from the lane's own rows, and the rows that cross
//     between the workgroup's waves are exchanged through LDS once per level.
//   * Each level's valid region shrinks by one cell per side; the ghost rim
//     absorbs it, and the last level stores only the item's interior.
//   * Domain faces apply the boundary rule at every level (clamp: the edge
//     cell itself, zero: 0), so partial results never leak across a face.
//   * After every pass of L steps the grid meets at a barrier (agent-scope
//     release/acquire, so stores from one XCD's L2 reach the others) and the
//     two field buffers swap roles.
//
// Arithmetic is identical to the per-step kernel (same order, no FMA
// contraction), so the result is bit-identical to the CPU gold.
#include "j3d_persistent.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "api_common.h"
#include "planner.h"

namespace perks {

namespace {

constexpr int kMaxLevels = 3;  // time steps per pass the kernel is built for

struct PersistArgs {
    int nz, ny, nx;
    int steps;
    int levels;                 // time steps per pass (L of the full passes)
    int nzs[kMaxLevels + 1];    // z-slabs per pass with l levels
    unsigned int* barrier;      // zeroed before the launch
};

// Grid-wide barrier for a cooperative launch.  Each wave releases its own
// stores at agent scope (s_waitcnt + L2 write-back), the workgroup meets, one
// lane counts the arrival, spins until every workgroup of this epoch has
// arrived, and every wave acquires (L2 invalidate) before reading again.
__device__ __forceinline__ void grid_barrier(unsigned int* counter, unsigned int nblocks, unsigned int epoch)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned int target = epoch * nblocks;
        while ((int)(__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) - target) < 0)
            __builtin_amdgcn_s_sleep(2);
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

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
    T* xch;                           // LDS exchange rows [2][kMaxLevels][NW][2][64]
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
        T* x = xch + ((size_t)(par * kMaxLevels + lv) * NW) * 2 * 64;
        x[(wave * 2 + 0) * 64 + lane] = out[0];
        x[(wave * 2 + 1) * 64 + lane] = out[NY - 1];
        __syncthreads();
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
        __syncthreads();
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
    __shared__ T xch[2 * kMaxLevels * NW * 2 * 64];
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
        default: run_pass<T, BC, NY, NW, P, 3, HAS_DIV>(src, dst, xch, a, k, lane, wave); break;
        }
        T* t = src;
        src = dst;
        dst = t;
        if (pass + 1 < npass) grid_barrier(a.barrier, gridDim.x, (unsigned int)(pass + 1));
    }
}

__global__ void zero_u32(unsigned int* p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Tile shapes the persistent kernel is built for: rows per wave, waves per
// workgroup, level-0 ring slots.  Variant 0 is the default; the others are
// selectable through opts->tune[4] for tuning sweeps.
struct Variant {
    int ny, nw, p;
};
constexpr Variant kVariants[] = {{4, 8, 6}, {8, 4, 3}, {4, 4, 6}, {2, 16, 6}};
constexpr int kNumVariants = 4;

template <typename T, int BC, bool HAS_DIV, int V>
void* kernel_ptr()
{
    constexpr Variant v = kVariants[V];
    return reinterpret_cast<void*>(&j3d_star1_persistent_kernel<T, BC, v.ny, v.nw, v.p, HAS_DIV>);
}

template <typename T, int BC, bool HAS_DIV>
void* kernel_ptr_v(int var)
{
    switch (var) {
    case 1: return kernel_ptr<T, BC, HAS_DIV, 1>();
    case 2: return kernel_ptr<T, BC, HAS_DIV, 2>();
    case 3: return kernel_ptr<T, BC, HAS_DIV, 3>();
    default: return kernel_ptr<T, BC, HAS_DIV, 0>();
    }
}

template <typename T>
void* pick_kernel(int bc, bool has_div, int var)
{
    if (bc == PERKS_BC_ZERO)
        return has_div ? kernel_ptr_v<T, PERKS_BC_ZERO, true>(var) : kernel_ptr_v<T, PERKS_BC_ZERO, false>(var);
    return has_div ? kernel_ptr_v<T, PERKS_BC_CLAMP, true>(var) : kernel_ptr_v<T, PERKS_BC_CLAMP, false>(var);
}

int variant_of(const perks_stencil_opts* o)
{
    const int v = o->tune[4];
    return (v >= 0 && v < kNumVariants) ? v : 0;
}

// Device scratch for the grid barrier, one per device, kept until perks_release().
std::mutex g_scratch_mu;
unsigned int* g_barrier[64] = {};

int barrier_scratch(unsigned int** out)
{
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return PERKS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_barrier[dev]) PERKS_HIP_TRY(hipMalloc(&g_barrier[dev], 256));
    *out = g_barrier[dev];
    return PERKS_OK;
}

}  // namespace

// ---- planning --------------------------------------------------------------

// Makespan model of one pass with `l` levels: items are dealt round-robin to
// G workgroups; an item costs its slab height plus the 2*l ghost planes it
// reads beyond the slab.  Picks the z-slab count that minimises
// rounds * (planes per item).
static int best_nzs(int nz, int ny, int nx, int l, int G, int rows)
{
    const int VW = 64 - 2 * l;
    const int VH = rows - 2 * (l - 1);
    const long long tiles = (long long)((nx + VW - 1) / VW) * ((ny + VH - 1) / VH);
    int best = 1;
    double best_cost = 1e300;
    const int cap = nz < 256 ? nz : 256;
    for (int s = 1; s <= cap; ++s) {
        const long long items = tiles * s;
        const long long rounds = (items + G - 1) / G;
        const double per = (double)(nz + s - 1) / s + 2.0 * l;
        const double cost = (double)rounds * per;
        if (cost < best_cost - 1e-9) {
            best_cost = cost;
            best = s;
        }
    }
    return best;
}

int plan_persistent(int dtype, int nz, int ny, int nx, const perks_stencil_opts* o, const perks_device_caps* caps,
                    int occupancy, perks_plan_info* out)
{
    if (o->shape != PERKS_SHAPE_3D_STAR_R1) return PERKS_ERR_UNSUPPORTED;
    if ((long long)ny * nx * (dtype == PERKS_F64 ? 8 : 4) >= 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
    if (occupancy <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
    int occ = occupancy;
    if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
    if (o->tune[3] > 0 && o->tune[3] < occ) occ = o->tune[3];
    const int G = caps->cu_count * occ;
    if (G <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
    int L = o->tune[2] > 0 ? o->tune[2] : 3;
    if (L > kMaxLevels) L = kMaxLevels;
    std::memset(out, 0, sizeof(*out));
    out->grid[0] = G;
    out->grid[1] = 1;
    out->grid[2] = 1;
    const Variant v = kVariants[variant_of(o)];
    out->block = v.nw * 64;
    out->occupancy = occ;
    out->lds_bytes = (int)(2 * kMaxLevels * v.nw * 2 * 64 * (dtype == PERKS_F64 ? 8 : 4));
    out->tile_x = 64;
    out->tile_y = v.nw * v.ny;
    out->reserved[0] = L;
    for (int l = 1; l <= kMaxLevels; ++l) out->reserved[l] = best_nzs(nz, ny, nx, l, G, v.nw * v.ny);
    return PERKS_OK;
}

int min_height_persistent(int, int, int, const perks_stencil_opts*) { return PERKS_ERR_UNSUPPORTED; }
int min_width_y_persistent(int, int, int, int, int) { return PERKS_ERR_UNSUPPORTED; }

// ---- launch -----------------------------------------------------------------

template <typename T>
int run_persistent(T* d_a, T* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf, const perks_stencil_opts* o,
                   hipStream_t stream, perks_run_report* rep)
{
    if (o->shape != PERKS_SHAPE_3D_STAR_R1) return PERKS_ERR_UNSUPPORTED;
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    if (!caps.cooperative) return PERKS_ERR_NOT_LAUNCHABLE;
    const bool has_div = cf.divisor != 0.0;
    const int var = variant_of(o);
    void* fn = pick_kernel<T>(o->boundary, has_div, var);
    int occ = 0;
    PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(fn),
                                                               kVariants[var].nw * 64, 0));
    perks_plan_info pl;
    const int dtype = sizeof(T) == 8 ? PERKS_F64 : PERKS_F32;
    rc = plan_persistent(dtype, nz, ny, nx, o, &caps, occ, &pl);
    if (rc != PERKS_OK) return rc;

    unsigned int* bar = nullptr;
    rc = barrier_scratch(&bar);
    if (rc != PERKS_OK) return rc;

    PersistArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nz = nz;
    a.ny = ny;
    a.nx = nx;
    a.steps = steps;
    a.levels = pl.reserved[0];
    for (int l = 1; l <= kMaxLevels; ++l) a.nzs[l] = pl.reserved[l];
    a.barrier = bar;
    KCoefs kc = cf;
    void* args[] = {&d_a, &d_b, &a, &kc};

    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    hipLaunchKernelGGL(zero_u32, dim3(1), dim3(64), 0, stream, bar, 64);
    PERKS_HIP_TRY(hipGetLastError());
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    if (steps > 0)
        PERKS_HIP_TRY(hipLaunchCooperativeKernel(fn, dim3(pl.grid[0]), dim3(pl.block), args, 0, stream));
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (rep) {
        const int npass = steps > 0 ? (steps + a.levels - 1) / a.levels : 0;
        rep->elapsed_ms = ms;
        rep->gcells_per_s = ms > 0 ? (double)steps * nz * ny * (double)nx / ms / 1e6 : 0.0;
        rep->grid[0] = pl.grid[0];
        rep->grid[1] = 1;
        rep->grid[2] = 1;
        rep->block = pl.block;
        rep->lds_bytes = pl.lds_bytes;
        rep->occupancy = pl.occupancy;
        rep->result_in_b = npass % 2;
        rep->reserved[0] = a.levels;
        rep->reserved[1] = npass;
    }
    return PERKS_OK;
}

template int run_persistent<double>(double*, double*, int, int, int, int, const KCoefs&, const perks_stencil_opts*,
                                    hipStream_t, perks_run_report*);
template int run_persistent<float>(float*, float*, int, int, int, int, const KCoefs&, const perks_stencil_opts*,
                                   hipStream_t, perks_run_report*);

int release_persistent_scratch()
{
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev >= 0 && dev < 64 && g_barrier[dev]) {
        PERKS_HIP_TRY(hipFree(g_barrier[dev]));
        g_barrier[dev] = nullptr;
    }
    return PERKS_OK;
}

}  // namespace perks
