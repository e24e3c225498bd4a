// j2d.hip -- 2D Jacobi kernels (all 2D shapes, FP64 / FP32 / FP16- and BF16-storage).
//
// Replaces the reference's 2D paths: the per-step kernel_baseline(_box)
// (stencil/2dstencil/jacobi-baseline.cu:28-175, launched per step from
// jacobi-2d.cu:623-630) and the persistent inner_general / kernel_general_wrapper
// (jacobi-general-kernel.cuh:23-587, jacobi-general-wrapper.cu:14-32, cooperative
// launch jacobi-2d.cu:754-763).
//
// One kernel serves both.  A workgroup stages a TX x TY tile of the field,
// including a ghost rim of R*L cells (R = shape radius, L = time steps in the
// pass), in LDS, runs the L steps in LDS (ping-pong between two LDS tiles, one
// workgroup barrier per step), and writes back the tile's interior.  With L = 1
// and one launch per step this is the plain per-step kernel; the persistent
// form is one cooperative launch in which every workgroup loops over passes and
// meets the grid at a barrier between them, so each pass reads and writes the
// field once for L steps.
//
// The work list can span several independent grids (the batched form): each
// grid has its own size, buffers and step count; grids are sorted by pass
// count so the grids still active in a pass are a prefix of the list.
//
// The kernel template itself is in j2d_kernel.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "api_common.h"
#include "j2d.h"

namespace perks {

namespace {

void* kernel_for(int dtype, int shape, int bc, bool hd)
{
    switch (dtype) {
    case PERKS_F64: return j2d_kernel_f64(shape, bc, hd);
    case PERKS_F32: return j2d_kernel_f32(shape, bc, hd);
    case PERKS_F16: return j2d_kernel_f16(shape, bc, hd);
    case PERKS_BF16: return j2d_kernel_bf16(shape, bc, hd);
    default: return nullptr;
    }
}

// Tile sizes of the instantiated kernels (j2d_kernel.h TileOf): 64 x 64 for
// FP64, 128 x 64 for the FP32-arithmetic types; two LDS tiles of 32 KiB.
void tile_dims(int dtype, int* tx, int* ty)
{
    *tx = dtype == PERKS_F64 ? 64 : 128;
    *ty = 64;
}

std::mutex g_mu2;
J2dGrid* g_grids[64] = {};
int* g_active[64] = {};
size_t g_grids_cap[64] = {}, g_active_cap[64] = {};

int scratch2(int ngrids, int npass, J2dGrid** grids, int** active)
{
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return PERKS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_mu2);
    if (g_grids_cap[dev] < (size_t)ngrids) {
        if (g_grids[dev]) PERKS_HIP_TRY(hipFree(g_grids[dev]));
        PERKS_HIP_TRY(hipMalloc(&g_grids[dev], sizeof(J2dGrid) * (size_t)ngrids));
        g_grids_cap[dev] = (size_t)ngrids;
    }
    if (g_active_cap[dev] < (size_t)npass) {
        if (g_active[dev]) PERKS_HIP_TRY(hipFree(g_active[dev]));
        PERKS_HIP_TRY(hipMalloc(&g_active[dev], sizeof(int) * (size_t)npass));
        g_active_cap[dev] = (size_t)npass;
    }
    *grids = g_grids[dev];
    *active = g_active[dev];
    return PERKS_OK;
}

}  // namespace

int j2d_pass_levels(int dtype, int shape, const perks_stencil_opts* o)
{
    const ShapeDesc* sd = shape_by_id(shape);
    if (!sd) return 1;
    int tx, ty;
    tile_dims(dtype, &tx, &ty);
    if (o->method == PERKS_METHOD_PLAIN) return 1;
    int L = o->tune[2] > 0 ? o->tune[2] : 0;
    if (L <= 0) {
        // Default: ghost rims of about a quarter of the shorter tile side.
        L = std::max(1, (std::min(tx, ty) / 4) / sd->radius);
        L = std::min(L, 8);
    }
    const int maxL = (std::min(tx, ty) - 2) / (2 * sd->radius);  // at least 2 interior cells per side
    return std::max(1, std::min(L, maxL));
}

int j2d_run(int dtype, J2dSpec* specs, int n, const perks_stencil_opts* o, hipStream_t stream,
            perks_run_report* rep)
{
    if (!o || !specs || n <= 0) return PERKS_ERR_INVALID_ARG;
    const ShapeDesc* sd = shape_by_id(o->shape);
    if (!sd || sd->dims != 2) return PERKS_ERR_INVALID_ARG;
    if (o->boundary != PERKS_BC_CLAMP && o->boundary != PERKS_BC_ZERO) return PERKS_ERR_INVALID_ARG;
    if (o->rounding != PERKS_ROUND_NEAREST && o->rounding != PERKS_ROUND_STOCHASTIC) return PERKS_ERR_INVALID_ARG;
    KCoefs cf;
    int rc = resolve_coefs(o, &cf, nullptr);
    if (rc != PERKS_OK) return rc;
    perks_device_caps caps;
    rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;

    int tx, ty;
    tile_dims(dtype, &tx, &ty);
    const int L = j2d_pass_levels(dtype, o->shape, o);
    const int VW = tx - 2 * sd->radius * L, VH = ty - 2 * sd->radius * L;

    // Sort by pass count (descending) so the grids active in a pass are a prefix.
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto npass_of = [&](int i) { return specs[i].steps > 0 ? (specs[i].steps + L - 1) / L : 0; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return npass_of(a) > npass_of(b); });
    std::vector<J2dGrid> grids;
    grids.reserve(n);
    int items = 0;
    long long cells_steps = 0;
    for (int i : order) {
        const J2dSpec& s = specs[i];
        if (s.ny <= 0 || s.nx <= 0 || s.steps < 0 || !s.a || !s.b) return PERKS_ERR_INVALID_ARG;
        if (s.steps == 0) continue;
        // Tiles are laid out with the full-pass geometry; a shorter last pass
        // reuses them (a smaller ghost rim is enough for it).
        const int ntx = (s.nx + VW - 1) / VW, nty = (s.ny + VH - 1) / VH;
        J2dGrid g;
        g.a = s.a;
        g.b = s.b;
        g.ny = s.ny;
        g.nx = s.nx;
        g.steps = s.steps;
        g.tiles = ntx * nty;
        items += g.tiles;
        g.item_end = items;
        grids.push_back(g);
        cells_steps += (long long)s.steps * s.ny * s.nx;
    }
    for (int i = 0; i < n; ++i) specs[i].result_in_b = npass_of(i) % 2;
    if (grids.empty()) {
        if (rep) std::memset(rep, 0, sizeof(*rep));
        return PERKS_OK;
    }
    const int npass = (grids[0].steps + L - 1) / L;
    std::vector<int> active(npass);
    for (int p = 0; p < npass; ++p) {
        int c = 0;
        while (c < (int)grids.size() && (grids[c].steps + L - 1) / L > p) ++c;
        active[p] = c;
    }
    // item_end is cumulative over the sorted list, so a prefix of grids covers
    // a prefix of items.  Tiles per grid do not depend on the pass.
    J2dGrid* d_grids = nullptr;
    int* d_active = nullptr;
    rc = scratch2((int)grids.size(), npass, &d_grids, &d_active);
    if (rc != PERKS_OK) return rc;
    GridSync* gs = nullptr;
    rc = grid_sync_get(kSyncJ2d, &gs);
    if (rc != PERKS_OK) return rc;
    PERKS_HIP_TRY(hipMemcpyAsync(d_grids, grids.data(), sizeof(J2dGrid) * grids.size(), hipMemcpyHostToDevice,
                                 stream));
    PERKS_HIP_TRY(hipMemcpyAsync(d_active, active.data(), sizeof(int) * npass, hipMemcpyHostToDevice, stream));

    void* fn = kernel_for(dtype, o->shape, o->boundary, cf.divisor != 0.0);
    if (!fn) return PERKS_ERR_UNSUPPORTED;
    J2dArgs a;
    a.grids = d_grids;
    a.active = d_active;
    a.levels = L;
    a.sync = gs;
    a.timeout = grid_sync_timeout();
    a.sr = (dtype == PERKS_F16 || dtype == PERKS_BF16) && o->rounding == PERKS_ROUND_STOCHASTIC;
    a.seed = o->round_seed;

    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    int grid_launched = 0, occ_used = 0;
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    if (o->method == PERKS_METHOD_PLAIN) {
        // One launch per step; each launch is a single pass over every tile.
        const int steps = grids[0].steps;
        for (int s = 0; s < steps; ++s) {
            // Per-step launches use the single-pass geometry of pass s.
            a.npass = 1;
            a.active = d_active + s;
            a.step_offset = s;
            const int gi = active[s] - 1;
            grid_launched = grids[gi].item_end;
            void* args[] = {&a, &cf};
            PERKS_HIP_TRY(hipLaunchKernel(fn, dim3(grid_launched), dim3(256), args, 0, stream));
        }
    } else {
        int occ = 0;
        PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(fn), 256, 0));
        if (occ <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
        if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
        if (o->tune[3] > 0 && o->tune[3] < occ) occ = o->tune[3];
        occ_used = occ;
        grid_launched = std::min(caps.cu_count * occ, items);
        a.npass = npass;
        a.step_offset = 0;
        void* args[] = {&a, &cf};
        PERKS_HIP_TRY(hipLaunchCooperativeKernel(fn, dim3(grid_launched), dim3(256), args, 0, stream));
    }
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (o->method != PERKS_METHOD_PLAIN) {
        rc = grid_sync_check(gs, stream);
        if (rc != PERKS_OK) return rc;
    }
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->elapsed_ms = ms;
        rep->gcells_per_s = ms > 0 ? (double)cells_steps / ms / 1e6 : 0.0;
        rep->grid[0] = grid_launched;
        rep->grid[1] = rep->grid[2] = 1;
        rep->block = 256;
        rep->lds_bytes = 2 * tx * ty * (dtype == PERKS_F64 ? 8 : 4);
        rep->occupancy = occ_used;
        rep->result_in_b = specs[0].result_in_b;
        rep->reserved[0] = L;
        rep->reserved[1] = npass;
    }
    return PERKS_OK;
}

int release_j2d_scratch()
{
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu2);
    if (dev < 0 || dev >= 64) return PERKS_OK;
    if (g_grids[dev]) PERKS_HIP_TRY(hipFree(g_grids[dev]));
    if (g_active[dev]) PERKS_HIP_TRY(hipFree(g_active[dev]));
    g_grids[dev] = nullptr;
    g_active[dev] = nullptr;
    g_grids_cap[dev] = g_active_cap[dev] = 0;
    return PERKS_OK;
}

}  // namespace perks
