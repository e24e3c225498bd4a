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
//     including a ghost rim of L cells in x and L-1 in y) times a z-slab.
//   * A workgroup marches through its item's slab plane by plane.  Level 0 is
//     the field as read from HBM; level l (1..L) is the field after l steps.
//     At each plane, level l computes plane p-l from three planes of level l-1
//     that it holds in VGPRs (a ring of three plane buffers per level), so a
//     plane is read from HBM once and written once per L steps.
//   * East/west neighbours come from the adjacent lane by DPP whole-wave
//     shifts, north/south from the lane's own rows.  The rows that cross
//     between the workgroup's waves are published to LDS when a level
//     produces a plane and read by the level above one iteration later, so
//     the workgroup meets once per plane for all levels.
//   * Each level's valid region shrinks by one cell per side; the ghost rim
//     absorbs it, and the last level stores only the item's interior.
//   * Domain faces apply the boundary rule at every level (clamp: the edge
//     cell itself, zero: 0), so partial results never leak across a face:
//     lanes beyond an x face hold the boundary value, waves at a y face and
//     planes at a z face select it.
//   * After every pass of L steps the grid meets at a barrier (agent-scope
//     release/acquire, so stores from one XCD's L2 reach the others) and the
//     two field buffers swap roles.
//
// Arithmetic is identical to the per-step kernel (same order, no FMA
// contraction), so the result is bit-identical to the CPU gold.
#include "j3d_persistent.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "api_common.h"
#include "planner.h"
#include "j3d_persistent_kernel.h"
#include "j3d_resident.h"

namespace perks {

namespace {

using j3dp::kMaxLevels;
using j3dp::kNumVariants;
using j3dp::PersistArgs;
using j3dp::Variant;
using j3dp::variant_shape;

void* pick_kernel(int fmt, int bc, bool has_div, int var, bool sr, int up)
{
    switch (fmt) {
    case PERKS_F64: return j3d_persistent_kernel_f64(bc, has_div, var, sr, up);
    case PERKS_F32: return j3d_persistent_kernel_f32(bc, has_div, var, sr, up);
    case PERKS_F16: return j3d_persistent_kernel_f16(bc, has_div, var, sr, up);
    case PERKS_BF16: return j3d_persistent_kernel_bf16(bc, has_div, var, sr, up);
    default: return nullptr;
    }
}

int elem_bytes(int fmt) { return fmt == PERKS_F64 ? 8 : (fmt == PERKS_F32 ? 4 : 2); }

int variant_of(const perks_stencil_opts* o)
{
    const int v = o->tune[4];
    return (v >= 0 && v < kNumVariants) ? v : 0;
}

// Time steps per pass when opts.tune[2] leaves it open, measured at 512^3 x
// 1000 steps on MI355X (variant 0, GCells/s at 3 / 4 levels).  FP64 and FP32
// take 4 (zero boundary: FP64 701 / 741, FP32 1265 / 1445 with ghost lanes at
// both faces; more since the edge tiles start at the faces).  16-bit storage:
// where the x edge tiles start at the faces (zero boundary, wide grids) a
// fourth level saves a tile column as well as traffic and 4 is fastest (BF16
// 1340 / 1399, FP16 1414 / 1485, stochastic rounding BF16 1025 / 1134, FP16
// 747 / 754); with ghost lanes at both faces (clamp) FP16 (1416 / 1290; with
// stochastic rounding 731 / 657) and BF16 with nearest rounding (1320 / 1250)
// take 3, where the ghost rim's arithmetic outweighs the traffic saved.
int default_levels(int dtype, bool sr, int bc, int nx)
{
    if (dtype == PERKS_F64 || dtype == PERKS_F32) return kMaxLevels;
    if (j3dp::x_face_tiles(bc, nx, kMaxLevels)) return kMaxLevels;
    if (dtype == PERKS_F16) return 3;
    return sr ? kMaxLevels : 3;
}

// Columns a tile stores with l steps per pass: 64 - 2l, or for the
// two-column variant 128 minus a rim of l rounded up to even per side.
int tile_width(const Variant& v, int l) { return v.pair ? 128 - 4 * ((l + 1) / 2) : 64 - 2 * l; }

}  // namespace

// ---- planning --------------------------------------------------------------

// Makespan model of one pass with `l` levels: items are dealt round-robin to
// G workgroups; an item costs its slab height plus the 2*l ghost planes it
// reads beyond the slab.  Picks the z-slab count that minimises
// rounds * (planes per item).
static int best_nzs(int nz, int ny, int ntx, int l, int G, int rows)
{
    const long long tiles = (long long)ntx * j3dp::y_tiles(ny, rows, l);
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
    if ((long long)ny * nx * elem_bytes(dtype) >= (long long)kOob) return PERKS_ERR_UNSUPPORTED_SIZE;
    if (occupancy <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
    if (o->tune[4] == kResidentVariant) {
        int occ = std::min(occupancy, std::max(1, kCoopMaxWavesPerCu / 8));
        if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
        return plan_resident(dtype, nz, ny, nx, caps, occ, out);
    }
    const Variant v = variant_shape(dtype, variant_of(o));
    if (v.pair && (nx & 1)) return PERKS_ERR_UNSUPPORTED;  // two columns per lane need an even width
    int occ = std::min(occupancy, std::max(1, kCoopMaxWavesPerCu / v.nw));
    if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
    if (o->tune[3] > 0 && o->tune[3] < occ) occ = o->tune[3];
    const int G = caps->cu_count * occ;
    if (G <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
    int L = o->tune[2] > 0 ? o->tune[2]
                           : default_levels(dtype, o->rounding == PERKS_ROUND_STOCHASTIC, o->boundary, nx);
    if (L > kMaxLevels) L = kMaxLevels;
    std::memset(out, 0, sizeof(*out));
    out->grid[0] = G;
    out->grid[1] = 1;
    out->grid[2] = 1;
    out->block = v.nw * 64;
    out->occupancy = occ;
    out->lds_bytes = (int)(2 * (kMaxLevels - 1) * v.nw * 2 * 64 * (dtype == PERKS_F64 || v.pair ? 8 : 4));
    out->tile_x = v.pair ? 128 : 64;
    out->tile_y = v.nw * v.ny;
    out->reserved[0] = L;
    for (int l = 1; l <= kMaxLevels; ++l)
    {
        // x tiles as the kernel cuts them (j3dp::x_tiles; the two-column form
        // keeps ghost lanes at both faces)
        const int ntx = v.pair ? (nx + tile_width(v, l) - 1) / tile_width(v, l) : j3dp::x_tiles(o->boundary, nx, l);
        out->reserved[l] = o->tune[7] > 0 ? std::min(o->tune[7], nz) : best_nzs(nz, ny, ntx, l, G, v.nw * v.ny);
    }
    return PERKS_OK;
}

// The reference's getminHeight (j3d.cu:598,612) is the domain height at which
// its persistent grid's register and shared-memory caches hold every plane.
// For the 7-point star in FP64/FP32 that is the height the resident-tile
// kernel (j3d_resident.hip) holds: as many layers of its blocks as the grid's
// waves cover.  Other shapes and 16-bit storage stream planes and keep only
// ring buffers on chip; for them the figure is the height whose field fits
// the chip's aggregate on-chip storage (LDS plus VGPRs of every CU), the size
// below which a persistent kernel could keep the whole domain resident.
int min_height_caps(int dtype, int ny, int nx, const perks_device_caps* caps)
{
    if (ny <= 0 || nx <= 0 || !caps) return PERKS_ERR_INVALID_ARG;
    const long long onchip = (long long)caps->cu_count *
                             ((long long)caps->lds_per_cu +
                              (long long)caps->vgprs_per_simd * caps->simds_per_cu * caps->warp_size * 4LL);
    const long long plane = (long long)ny * nx * elem_bytes(dtype);
    const long long h = onchip / plane;
    if (h < 3) return PERKS_ERR_HEIGHT_TOO_SMALL;
    return (int)std::min<long long>(h, 0x7fffffffLL);
}

int min_height_persistent(int dtype, int ny, int nx, const perks_stencil_opts* o)
{
    perks_device_caps caps;
    const int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    const bool star = !o || (o->shape == PERKS_SHAPE_3D_STAR_R1 && o->tune[5] != 1);
    if (star && (dtype == PERKS_F64 || dtype == PERKS_F32)) {
        if (ny <= 0 || nx <= 0) return PERKS_ERR_INVALID_ARG;
        const int bc = o ? o->boundary : PERKS_BC_CLAMP;
        int occ = 0;
        const int r = resident_occupancy(dtype, bc, false, &occ);
        if (r != PERKS_OK) return r;
        occ = std::min(occ, std::max(1, kCoopMaxWavesPerCu / 8));
        if (o && o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
        return resident_height(dtype, ny, nx, &caps, occ);
    }
    return min_height_caps(dtype, ny, nx, &caps);
}

// getMinWidthY (j3d.cu:69-123) is the largest of getminHeight over the
// workgroups-per-CU and cache variants; the on-chip capacity above does not
// depend on them.
int min_width_y_persistent(int dtype, int nx, int ny, int, int)
{
    return min_height_persistent(dtype, ny, nx, nullptr);
}

int j3d_persistent_variants() { return kNumVariants; }

int j3d_persistent_update(int dtype, const perks_stencil_opts* o, const KCoefs& cf)
{
    if (o->tune[6] == 1) return j3dp::kUpdFused;
    bool same = false;
    if (dtype == PERKS_F64) {
        const double k = cf.c[0];
        same = k > 0.0;
        for (int i = 1; i < 6; ++i) same = same && cf.c[i] == k;
    } else {  // FP32 arithmetic: FP32, FP16 and BF16 storage
        const float k = (float)cf.c[0];
        same = k > 0.0f;
        for (int i = 1; i < 6; ++i) same = same && (float)cf.c[i] == k;
    }
    return same ? j3dp::kUpdShared : j3dp::kUpdGold;
}

void* j3d_persistent_fn(int dtype, int bc, bool has_div, int var, bool sr, int up, int* block)
{
    if (var < 0 || var >= kNumVariants) return nullptr;
    if (block) *block = variant_shape(dtype, var).nw * 64;
    return pick_kernel(dtype, bc, has_div, var, sr, up);
}

// ---- launch -----------------------------------------------------------------

namespace {

int run_persistent_fmt(int dtype, void* d_a, void* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
                       const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep)
{
    if (o->shape != PERKS_SHAPE_3D_STAR_R1) return PERKS_ERR_UNSUPPORTED;
    if (o->tune[4] == kResidentVariant) {
        if (o->tune[6] == 1) return PERKS_ERR_UNSUPPORTED;  // the contracted update is built for the ring kernel only
        return run_resident(dtype, d_a, d_b, nz, ny, nx, steps, cf, o, stream, rep);
    }
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    if (!caps.cooperative) return PERKS_ERR_NOT_LAUNCHABLE;
    const bool has_div = cf.divisor != 0.0;
    const int var = variant_of(o);
    // tune[6] = 1: the contracted update (FP64/FP32 without a divisor; the
    // kernel pickers return nullptr for the rest)
    void* fn = pick_kernel(dtype, o->boundary, has_div, var, o->rounding == PERKS_ROUND_STOCHASTIC,
                           j3d_persistent_update(dtype, o, cf));
    if (!fn) return PERKS_ERR_UNSUPPORTED;
    int occ = 0;
    PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(fn),
                                                               variant_shape(dtype, var).nw * 64, 0));
    perks_plan_info pl;
    rc = plan_persistent(dtype, nz, ny, nx, o, &caps, occ, &pl);
    if (rc != PERKS_OK) return rc;

    SyncLease lease;
    rc = lease.acquire();
    if (rc != PERKS_OK) return rc;
    GridSync* gs = lease.gs;

    PersistArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nz = nz;
    a.ny = ny;
    a.nx = nx;
    a.steps = steps;
    a.levels = pl.reserved[0];
    for (int l = 1; l <= kMaxLevels; ++l) a.nzs[l] = pl.reserved[l];
    a.sync = gs;
    a.timeout = grid_sync_timeout();
    a.seed = o->round_seed;
    KCoefs kc = cf;
    void* args[] = {&d_a, &d_b, &a, &kc};

    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    if (steps > 0)
        PERKS_HIP_TRY(launch_persistent(fn, dim3(pl.grid[0]), dim3(pl.block), args, 0, stream));
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    rc = grid_sync_check(gs, stream);
    if (rc != PERKS_OK) return rc;
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

}  // namespace

template <typename T>
int run_persistent(T* d_a, T* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf, const perks_stencil_opts* o,
                   hipStream_t stream, perks_run_report* rep)
{
    return run_persistent_fmt(sizeof(T) == 8 ? PERKS_F64 : PERKS_F32, d_a, d_b, nz, ny, nx, steps, cf, o, stream,
                              rep);
}

template int run_persistent<double>(double*, double*, int, int, int, int, const KCoefs&, const perks_stencil_opts*,
                                    hipStream_t, perks_run_report*);
template int run_persistent<float>(float*, float*, int, int, int, int, const KCoefs&, const perks_stencil_opts*,
                                   hipStream_t, perks_run_report*);

int run_persistent16(int fmt, uint16_t* d_a, uint16_t* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
                     const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep)
{
    if (fmt != PERKS_F16 && fmt != PERKS_BF16) return PERKS_ERR_INVALID_ARG;
    return run_persistent_fmt(fmt, d_a, d_b, nz, ny, nx, steps, cf, o, stream, rep);
}

int release_persistent_scratch()
{
    return grid_sync_release();
}

}  // namespace perks
