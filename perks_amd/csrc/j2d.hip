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
// count so the grids still active in a pass are a prefix of the list.  A batch
// plan (perks_jacobi2d_batch_plan) uploads the work list once, picks the tile
// shape, the steps per pass and the workgroups per CU by timing candidates on
// scratch buffers of the batch's sizes, and captures the launch in a hipGraph
// that later calls replay.
//
// Tile variants 2 and 4 are a different kernel (j2d_stream_kernel.h): each
// wave streams the rows of a strip through register rings, one per time
// level, with no LDS; its work items are strips times row segments.  Variant
// 2 holds one column per lane (64-column strips), variant 4 two (128-column
// strips, FP32/FP16/BF16 storage and even grid widths); the default is 4
// where it applies, else 2.
//
// The LDS kernel template is in j2d_kernel.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "api_common.h"
#include "j2d.h"

namespace perks {

namespace {

constexpr int kStreamTile = 2;     // tile variant of the row-streaming kernel, one column per lane
constexpr int kPairTile = 4;       // the row-streaming kernel, two columns per lane
constexpr int kDefaultTile = -1;   // resolved per run: kPairTile where it applies, else kStreamTile
constexpr int kStreamWaves = 4;    // its waves per workgroup, one work item each
constexpr int kStreamItems = 16384;  // work items a launch aims for (segments are cut to reach it)

void* kernel_for(int dtype, int shape, int bc, bool hd, int tile)
{
    if (tile == kStreamTile || tile == kPairTile) {
        const bool pair = tile == kPairTile;
        switch (dtype) {
        case PERKS_F64: return j2d_stream_kernel_f64(shape, bc, hd, pair);
        case PERKS_F32: return j2d_stream_kernel_f32(shape, bc, hd, pair);
        case PERKS_F16: return j2d_stream_kernel_f16(shape, bc, hd, pair);
        case PERKS_BF16: return j2d_stream_kernel_bf16(shape, bc, hd, pair);
        default: return nullptr;
        }
    }
    switch (dtype) {
    case PERKS_F64: return j2d_kernel_f64(shape, bc, hd, tile);
    case PERKS_F32: return j2d_kernel_f32(shape, bc, hd, tile);
    case PERKS_F16: return j2d_kernel_f16(shape, bc, hd, tile);
    case PERKS_BF16: return j2d_kernel_bf16(shape, bc, hd, tile);
    default: return nullptr;
    }
}

// Tile sizes of the instantiated LDS kernels (j2d_kernel.h TileOf); the
// streaming kernel's strips are 64 columns wide and its segments are cut per
// work list.
void tile_dims(int dtype, int tile, int* tx, int* ty)
{
    if (tile == kStreamTile || tile == kPairTile) {
        *tx = tile == kPairTile ? 128 : 64;
        *ty = 0;
        return;
    }
    *tx = dtype == PERKS_F64 ? 64 : 128;
    *ty = tile == 1 ? 64 : 32;
}

int max_levels(int dtype, int tile, int radius)
{
    if (tile == kStreamTile)  // registers, and at least 8 stored columns per strip
        return std::max(1, std::min(stream_max_levels(radius), (64 - 8) / (2 * radius)));
    if (tile == kPairTile) {
        int l = stream_max_levels(radius);
        while (l > 1 && 128 - 2 * pair_ghost(radius, l) < 8) --l;
        return l;
    }
    int tx, ty;
    tile_dims(dtype, tile, &tx, &ty);
    return std::max(1, (std::min(tx, ty) - 2) / (2 * radius));  // at least 2 interior cells per side
}

// opts->tune[4]: 0 = default, 1 = LDS tile 1, 2 = streaming (one column per
// lane), 3 = LDS tile 0, 4 = streaming (two columns per lane).  Returned as
// the internal variant (0, 1: LDS tiles, 2, 4: streaming; kDefaultTile).
int tile_of(const perks_stencil_opts* o)
{
    switch (o->tune[4]) {
    case 1: return 1;
    case 2: return kStreamTile;
    case 3: return 0;
    case 4: return kPairTile;
    default: return kDefaultTile;
    }
}

// The two-column kernel needs FP32/FP16/BF16 storage and even widths.
bool pair_applies(int dtype, const J2dSpec* specs, int n)
{
    if (dtype == PERKS_F64) return false;
    for (int i = 0; i < n; ++i)
        if (specs[i].steps > 0 && (specs[i].nx & 1)) return false;
    return true;
}

// The variant a run uses: the requested one, the default resolved, or an
// error when the two-column kernel was asked for where it does not apply.
int resolve_tile(int dtype, const perks_stencil_opts* o, const J2dSpec* specs, int n, int* tile)
{
    int t = tile_of(o);
    const bool pair_ok = pair_applies(dtype, specs, n);
    if (t == kDefaultTile) t = pair_ok ? kPairTile : kStreamTile;
    if (t == kPairTile && !pair_ok) return PERKS_ERR_UNSUPPORTED;
    *tile = t;
    return PERKS_OK;
}

// Workgroups a launch of `items` work items needs (one item per LDS
// workgroup, one per wave of the streaming kernel).
int groups_for(int tile, int items)
{
    return (tile == kStreamTile || tile == kPairTile) ? (items + kStreamWaves - 1) / kStreamWaves : items;
}

// A launch configuration: tile variant, time steps per pass, workgroups per
// CU (0: as many as fit).
struct Config {
    int tile;
    int levels;
    int occ;
};

// The host-side work list of a batch for a given pass depth.
struct Work {
    std::vector<J2dGrid> grids;  // sorted by pass count (descending); item_end cumulative
    std::vector<int> active;     // grids still running in each pass (a prefix of `grids`)
    std::vector<int> result_in_b;  // per input grid
    int npass = 0;
    int items = 0;
    long long cell_steps = 0;
};

int build_work(const J2dSpec* specs, int n, int R, int tx, int ty, int L, Work* w)
{
    const bool stream = ty == 0;
    const int VW = (stream && tx == 128) ? 128 - 2 * pair_ghost(R, L) : tx - 2 * R * L;
    int VH = ty - 2 * R * L;
    if (VW < 1 || (!stream && VH < 1)) return PERKS_ERR_INVALID_ARG;
    if (stream) {
        // Segment height: enough work items to fill the chip, and at least
        // 8*R*L rows so that the ghost rows stay a small part of an item.
        long long strip_rows = 0;
        for (int i = 0; i < n; ++i)
            if (specs[i].steps > 0 && specs[i].nx > 0 && specs[i].ny > 0)
                strip_rows += (long long)((specs[i].nx + VW - 1) / VW) * specs[i].ny;
        const long long seg = std::max<long long>(8LL * R * L, (strip_rows + kStreamItems - 1) / kStreamItems);
        VH = (int)std::min<long long>(seg, 0x7fffffffLL);
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto npass_of = [&](int i) { return specs[i].steps > 0 ? (specs[i].steps + L - 1) / L : 0; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return npass_of(a) > npass_of(b); });
    w->grids.clear();
    w->result_in_b.assign(n, 0);
    w->items = 0;
    w->cell_steps = 0;
    for (int i : order) {
        const J2dSpec& s = specs[i];
        if (s.ny <= 0 || s.nx <= 0 || s.steps < 0 || !s.a || !s.b) return PERKS_ERR_INVALID_ARG;
        w->result_in_b[i] = npass_of(i) % 2;
        if (s.steps == 0) continue;
        // Tiles are laid out with the full-pass geometry; a shorter last pass
        // reuses them (a smaller ghost rim is enough for it).
        const long long tiles = (long long)((s.nx + VW - 1) / VW) * ((s.ny + VH - 1) / VH);
        if (w->items + tiles > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
        J2dGrid g;
        g.a = s.a;
        g.b = s.b;
        g.ny = s.ny;
        g.nx = s.nx;
        g.steps = s.steps;
        g.tiles = (int)tiles;
        g.seg = stream ? VH : 0;
        w->items += g.tiles;
        g.item_end = w->items;
        w->grids.push_back(g);
        w->cell_steps += (long long)s.steps * s.ny * s.nx;
    }
    w->npass = w->grids.empty() ? 0 : (w->grids[0].steps + L - 1) / L;
    w->active.assign(w->npass, 0);
    for (int p = 0; p < w->npass; ++p) {
        int c = 0;
        while (c < (int)w->grids.size() && (w->grids[c].steps + L - 1) / L > p) ++c;
        w->active[p] = c;
    }
    return PERKS_OK;
}

// Device copies of a work list.
struct DeviceWork {
    J2dGrid* grids = nullptr;
    int* active = nullptr;
    size_t grids_cap = 0, active_cap = 0;

    int upload(const Work& w, hipStream_t stream)
    {
        if (grids_cap < w.grids.size()) {
            if (grids) PERKS_HIP_TRY(hipFree(grids));
            PERKS_HIP_TRY(hipMalloc(&grids, sizeof(J2dGrid) * w.grids.size()));
            grids_cap = w.grids.size();
        }
        if (active_cap < w.active.size()) {
            if (active) PERKS_HIP_TRY(hipFree(active));
            PERKS_HIP_TRY(hipMalloc(&active, sizeof(int) * w.active.size()));
            active_cap = w.active.size();
        }
        PERKS_HIP_TRY(hipMemcpyAsync(grids, w.grids.data(), sizeof(J2dGrid) * w.grids.size(), hipMemcpyHostToDevice,
                                     stream));
        PERKS_HIP_TRY(hipMemcpyAsync(active, w.active.data(), sizeof(int) * w.active.size(), hipMemcpyHostToDevice,
                                     stream));
        return PERKS_OK;
    }
    int release()
    {
        if (grids) PERKS_HIP_TRY(hipFree(grids));
        if (active) PERKS_HIP_TRY(hipFree(active));
        grids = nullptr;
        active = nullptr;
        grids_cap = active_cap = 0;
        return PERKS_OK;
    }
};

std::mutex g_mu2;
DeviceWork g_work[64];  // per device, for the direct (unplanned) calls

// Everything a launch needs once the work list is on the device.
struct Launch {
    void* fn = nullptr;
    int tile = 0;
    int grid = 0;     // persistent: workgroups; plain: unused (per-step grids from the work list)
    int occ = 0;
    J2dArgs args;
    KCoefs cf;
};

int prepare_launch(int dtype, const perks_stencil_opts* o, const KCoefs& cf, const Config& c, const Work& w,
                   const DeviceWork& dw, GridSync* gs, Launch* out)
{
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    out->fn = kernel_for(dtype, o->shape, o->boundary, cf.divisor != 0.0, c.tile);
    if (!out->fn) return PERKS_ERR_UNSUPPORTED;
    out->tile = c.tile;
    std::memset(&out->args, 0, sizeof(out->args));
    out->args.grids = dw.grids;
    out->args.active = dw.active;
    out->args.levels = c.levels;
    out->args.sync = gs;
    out->args.timeout = grid_sync_timeout();
    out->args.sr = (dtype == PERKS_F16 || dtype == PERKS_BF16) && o->rounding == PERKS_ROUND_STOCHASTIC;
    out->args.seed = o->round_seed;
    out->cf = cf;
    out->occ = 0;
    out->grid = 0;
    if (o->method != PERKS_METHOD_PLAIN) {
        int occ = 0;
        PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(out->fn), 256, 0));
        if (occ <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
        occ = std::min(occ, kCoopMaxWavesPerCu / 4);  // 256-thread workgroups
        if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
        if (c.occ > 0 && c.occ < occ) occ = c.occ;
        out->occ = occ;
        out->grid = std::min(caps.cu_count * occ, groups_for(c.tile, w.items));
        out->args.npass = w.npass;
        out->args.step_offset = 0;
    }
    return PERKS_OK;
}

// Enqueues the launch(es) on `stream`.  cooperative: hipLaunchCooperativeKernel
// (checked co-residency); otherwise an ordinary launch of the same grid, which
// a stream capture can record (the grid is at most CUs x occupancy, and the
// barrier's bounded wait turns a grid that is not co-resident into
// PERKS_ERR_BARRIER_TIMEOUT instead of a hang).
int enqueue(Launch& l, const perks_stencil_opts* o, const Work& w, hipStream_t stream, bool cooperative)
{
    if (w.grids.empty()) return PERKS_OK;
    if (o->method == PERKS_METHOD_PLAIN) {
        // One launch per step; each launch is a single pass over every tile.
        const int steps = w.grids[0].steps;
        J2dArgs a = l.args;
        a.npass = 1;
        for (int s = 0; s < steps; ++s) {
            a.active = l.args.active + s;
            a.step_offset = s;
            const int g = groups_for(l.tile, w.grids[w.active[s] - 1].item_end);
            void* args[] = {&a, &l.cf};
            PERKS_HIP_TRY(hipLaunchKernel(l.fn, dim3(g), dim3(256), args, 0, stream));
        }
        return PERKS_OK;
    }
    void* args[] = {&l.args, &l.cf};
    if (cooperative) PERKS_HIP_TRY(launch_persistent(l.fn, dim3(l.grid), dim3(256), args, 0, stream));
    else PERKS_HIP_TRY(hipLaunchKernel(l.fn, dim3(l.grid), dim3(256), args, 0, stream));
    return PERKS_OK;
}

int check_opts(const perks_stencil_opts* o, const ShapeDesc** sd_out)
{
    const ShapeDesc* sd = shape_by_id(o->shape);
    if (!sd || sd->dims != 2) return PERKS_ERR_INVALID_ARG;
    if (o->boundary != PERKS_BC_CLAMP && o->boundary != PERKS_BC_ZERO) return PERKS_ERR_INVALID_ARG;
    if (o->method != PERKS_METHOD_PLAIN && o->method != PERKS_METHOD_PERSISTENT) return PERKS_ERR_INVALID_ARG;
    if (o->rounding != PERKS_ROUND_NEAREST && o->rounding != PERKS_ROUND_STOCHASTIC) return PERKS_ERR_INVALID_ARG;
    *sd_out = sd;
    return PERKS_OK;
}

void fill_report(perks_run_report* rep, float ms, const Work& w, const Launch& l, const Config& c, int dtype)
{
    if (!rep) return;
    int tx, ty;
    tile_dims(dtype, c.tile, &tx, &ty);
    std::memset(rep, 0, sizeof(*rep));
    rep->elapsed_ms = ms;
    rep->gcells_per_s = ms > 0 ? (double)w.cell_steps / ms / 1e6 : 0.0;
    rep->grid[0] = l.grid > 0 ? l.grid : (w.grids.empty() ? 0 : w.grids.back().item_end);
    rep->grid[1] = rep->grid[2] = 1;
    rep->block = 256;
    rep->lds_bytes = 2 * tx * ty * (dtype == PERKS_F64 ? 8 : 4);  // 0 for the streaming kernel (ty = 0)
    rep->occupancy = l.occ;
    rep->result_in_b = w.result_in_b.empty() ? 0 : w.result_in_b[0];
    rep->reserved[0] = c.levels;
    rep->reserved[1] = w.npass;
    rep->reserved[2] = c.tile;
}

}  // namespace

int j2d_pass_levels(int dtype, int shape, const perks_stencil_opts* o, int tile)
{
    const ShapeDesc* sd = shape_by_id(shape);
    if (!sd) return 1;
    if (o->method == PERKS_METHOD_PLAIN) return 1;
    // Default: 4 / R steps per pass (the best depths measured at 8192^2 for
    // radius 1, 2 and 3-6 on the default tile).
    const int L = o->tune[2] > 0 ? o->tune[2] : std::max(1, 4 / sd->radius);
    return std::max(1, std::min(L, max_levels(dtype, tile, sd->radius)));
}

int j2d_run(int dtype, J2dSpec* specs, int n, const perks_stencil_opts* o, hipStream_t stream,
            perks_run_report* rep)
{
    if (!o || !specs || n <= 0) return PERKS_ERR_INVALID_ARG;
    const ShapeDesc* sd = nullptr;
    int rc = check_opts(o, &sd);
    if (rc != PERKS_OK) return rc;
    KCoefs cf;
    rc = resolve_coefs(o, &cf, nullptr);
    if (rc != PERKS_OK) return rc;
    int tile = 0;
    rc = resolve_tile(dtype, o, specs, n, &tile);
    if (rc != PERKS_OK) return rc;
    const Config c{tile, j2d_pass_levels(dtype, o->shape, o, tile), o->tune[3]};
    int tx, ty;
    tile_dims(dtype, c.tile, &tx, &ty);
    Work w;
    rc = build_work(specs, n, sd->radius, tx, ty, c.levels, &w);
    if (rc != PERKS_OK) return rc;
    for (int i = 0; i < n; ++i) specs[i].result_in_b = w.result_in_b[i];
    if (w.grids.empty()) {
        if (rep) std::memset(rep, 0, sizeof(*rep));
        return PERKS_OK;
    }
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return PERKS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_mu2);
    rc = g_work[dev].upload(w, stream);
    if (rc != PERKS_OK) return rc;
    SyncLease lease;
    rc = lease.acquire();
    if (rc != PERKS_OK) return rc;
    GridSync* gs = lease.gs;
    Launch l;
    rc = prepare_launch(dtype, o, cf, c, w, g_work[dev], gs, &l);
    if (rc != PERKS_OK) return rc;

    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    rc = enqueue(l, o, w, stream, true);
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (rc != PERKS_OK) return rc;
    if (o->method != PERKS_METHOD_PLAIN) {
        rc = grid_sync_check(gs, stream);
        if (rc != PERKS_OK) return rc;
    }
    fill_report(rep, ms, w, l, c, dtype);
    return PERKS_OK;
}

// ---- the batch without device allocation -----------------------------------
// Callers that cannot allocate device memory (tight serving loops, harnesses
// that forbid hipMalloc) run a batch through perks_jacobi2d_batch_noalloc.
// Everything the launch needs beyond the caller's fields -- the grid barrier
// block, the work list and the per-pass count of active grids -- lives in
// `scratch`, a file-scope device array of fixed size that each GPU holds for
// the life of the process.  In host code the array's name is only a symbol:
// its device address on the current GPU comes from hipGetSymbolAddress (once
// per device) and only that address is passed to kernels and copies.
namespace {

constexpr size_t kNoAllocSyncOff = 0;
constexpr size_t kNoAllocGridsOff = (sizeof(GridSync) + 255) / 256 * 256;
constexpr size_t kNoAllocActiveOff =
    kNoAllocGridsOff + ((size_t)PERKS_NOALLOC_MAX_GRIDS * sizeof(J2dGrid) + 255) / 256 * 256;
constexpr size_t kNoAllocBytes = kNoAllocActiveOff + (size_t)PERKS_NOALLOC_MAX_PASSES * sizeof(int);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_noalloc_mu;                 // one batch at a time per process and device
unsigned char* g_noalloc_ptr[64] = {};   // device address of `scratch` per device

// Device address of `scratch` on the current device (looked up on the first
// call on that device).  g_noalloc_mu held.
int noalloc_scratch(int dev, unsigned char** out)
{
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            std::fprintf(stderr, "perks_amd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return PERKS_ERR_HIP;
        }
        g_noalloc_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_noalloc_ptr[dev];
    return PERKS_OK;
}

}  // namespace

int j2d_run_noalloc(int dtype, J2dSpec* specs, int n, const perks_stencil_opts* o, hipStream_t stream,
                    perks_run_report* rep)
{
    if (!o || !specs || n <= 0) return PERKS_ERR_INVALID_ARG;
    const ShapeDesc* sd = nullptr;
    int rc = check_opts(o, &sd);
    if (rc != PERKS_OK) return rc;
    KCoefs cf;
    rc = resolve_coefs(o, &cf, nullptr);
    if (rc != PERKS_OK) return rc;
    int tile = 0;
    rc = resolve_tile(dtype, o, specs, n, &tile);
    if (rc != PERKS_OK) return rc;
    const Config c{tile, j2d_pass_levels(dtype, o->shape, o, tile), o->tune[3]};
    int tx, ty;
    tile_dims(dtype, c.tile, &tx, &ty);
    Work w;
    rc = build_work(specs, n, sd->radius, tx, ty, c.levels, &w);
    if (rc != PERKS_OK) return rc;
    if (w.grids.size() > (size_t)PERKS_NOALLOC_MAX_GRIDS || w.active.size() > (size_t)PERKS_NOALLOC_MAX_PASSES) {
        std::fprintf(stderr,
                     "perks_amd: perks_jacobi2d_batch_noalloc: %zu grids and %zu passes exceed its fixed scratch "
                     "(at most %d grids and %d passes)\n",
                     w.grids.size(), w.active.size(), PERKS_NOALLOC_MAX_GRIDS, PERKS_NOALLOC_MAX_PASSES);
        return PERKS_ERR_UNSUPPORTED_SIZE;
    }
    for (int i = 0; i < n; ++i) specs[i].result_in_b = w.result_in_b[i];
    if (rep) std::memset(rep, 0, sizeof(*rep));
    if (w.grids.empty()) return PERKS_OK;
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return PERKS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    unsigned char* base = nullptr;
    rc = noalloc_scratch(dev, &base);
    if (rc != PERKS_OK) return rc;
    GridSync* gs = reinterpret_cast<GridSync*>(base + kNoAllocSyncOff);
    DeviceWork dw;  // a view of the scratch: never released
    dw.grids = reinterpret_cast<J2dGrid*>(base + kNoAllocGridsOff);
    dw.active = reinterpret_cast<int*>(base + kNoAllocActiveOff);
    // The array keeps the last call's contents: the barrier block must start
    // at zero, the work list is rewritten whole.
    PERKS_HIP_TRY(hipMemsetAsync(gs, 0, sizeof(GridSync), stream));
    PERKS_HIP_TRY(hipMemcpyAsync(dw.grids, w.grids.data(), sizeof(J2dGrid) * w.grids.size(), hipMemcpyHostToDevice,
                                 stream));
    PERKS_HIP_TRY(hipMemcpyAsync(dw.active, w.active.data(), sizeof(int) * w.active.size(), hipMemcpyHostToDevice,
                                 stream));
    Launch l;
    rc = prepare_launch(dtype, o, cf, c, w, dw, gs, &l);
    if (rc != PERKS_OK) return rc;
    rc = enqueue(l, o, w, stream, true);
    if (rc != PERKS_OK) return rc;
    // Wait before the scratch can be rewritten by the next call.
    if (o->method != PERKS_METHOD_PLAIN) {
        rc = grid_sync_check(gs, stream);
        if (rc != PERKS_OK) return rc;
    } else {
        PERKS_HIP_TRY(hipStreamSynchronize(stream));
    }
    fill_report(rep, 0.f, w, l, c, dtype);
    return PERKS_OK;
}

int release_j2d_scratch()
{
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu2);
    if (dev < 0 || dev >= 64) return PERKS_OK;
    return g_work[dev].release();
}

}  // namespace perks

// ---- batch plans ---------------------------------------------------------------
struct perks_batch_plan {
    int dtype = 0;
    int device = 0;
    perks_stencil_opts opts;
    perks::Config cfg{0, 1, 0};
    perks::Work work;
    perks::DeviceWork dw;
    perks::Launch launch;
    perks::GridSync* gs = nullptr;  // the plan's own barrier block (grid_sync_alloc)
    std::mutex run_mu;              // runs of one plan share that block: one at a time
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;        // persistent method: the kernel node marked cooperative
    hipGraphExec_t exec_plain = nullptr;  // the same graph as an ordinary launch (fallback)
    int tuned = 0;
    double tuned_gcells = 0.0;
};

namespace perks {
namespace {

// A launch captured from a stream loses its cooperative property: the kernel
// node of a captured hipLaunchCooperativeKernel reads its cooperative
// attribute as 0 (HIP 7.x, MI355X).  Marking the node cooperative makes a
// replay a cooperative launch again, so the runtime, not the caller,
// guarantees that the persistent grid is co-resident.
int mark_cooperative(hipGraph_t g)
{
    size_t n = 0;
    PERKS_HIP_TRY(hipGraphGetNodes(g, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    PERKS_HIP_TRY(hipGraphGetNodes(g, nodes.data(), &n));
    for (hipGraphNode_t node : nodes) {
        hipGraphNodeType t;
        PERKS_HIP_TRY(hipGraphNodeGetType(node, &t));
        if (t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeAttrValue v;
        std::memset(&v, 0, sizeof(v));
        v.cooperative = 1;
        PERKS_HIP_TRY(hipGraphKernelNodeSetAttribute(node, hipKernelNodeAttributeCooperative, &v));
    }
    return PERKS_OK;
}

// Times one configuration on scratch buffers: every grid runs min(steps, cap)
// steps.  Returns cells per second (0 when the configuration does not apply).
double time_config(int dtype, const perks_stencil_opts* o, const KCoefs& cf, const Config& c,
                   const std::vector<J2dSpec>& scratch, int R, int cap, hipStream_t stream, GridSync* gs)
{
    int tx, ty;
    tile_dims(dtype, c.tile, &tx, &ty);
    if (c.levels > max_levels(dtype, c.tile, R)) return 0.0;
    std::vector<J2dSpec> specs = scratch;
    for (auto& s : specs) s.steps = std::min(s.steps, cap);
    Work w;
    if (build_work(specs.data(), (int)specs.size(), R, tx, ty, c.levels, &w) != PERKS_OK || w.grids.empty())
        return 0.0;
    DeviceWork dw;
    double best = 0.0;
    do {
        if (dw.upload(w, stream) != PERKS_OK) break;
        Launch l;
        if (prepare_launch(dtype, o, cf, c, w, dw, gs, &l) != PERKS_OK) break;
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess) break;
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            break;
        }
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms up
            if (hipEventRecord(e0, stream) != hipSuccess) break;
            if (enqueue(l, o, w, stream, true) != PERKS_OK) break;
            if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) break;
            if (o->method != PERKS_METHOD_PLAIN && grid_sync_check(gs, stream) != PERKS_OK) break;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
            if (rep == 1 && ms > 0) best = (double)w.cell_steps / (ms * 1e-3);
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } while (0);
    (void)dw.release();
    return best;
}

constexpr size_t kTuneScratchBytes = 4ull << 30;  // scratch buffers the autotuner may allocate
constexpr int kTuneSteps = 48;                    // time steps per grid in a trial (multiple of 1,2,3,4,6,8,12,16)

int plan_batch(int dtype, const perks_grid2d* grids, int ngrids, const perks_stencil_opts* o, int autotune,
               hipStream_t stream, perks_batch_plan** out)
{
    const ShapeDesc* sd = nullptr;
    int rc = check_opts(o, &sd);
    if (rc != PERKS_OK) return rc;
    KCoefs cf;
    rc = resolve_coefs(o, &cf, nullptr);
    if (rc != PERKS_OK) return rc;
    std::vector<J2dSpec> specs((size_t)ngrids);
    for (int i = 0; i < ngrids; ++i) {
        if (!grids[i].a || !grids[i].b || grids[i].width_y <= 0 || grids[i].width_x <= 0 || grids[i].steps < 0)
            return PERKS_ERR_INVALID_ARG;
        specs[i] = J2dSpec{grids[i].a, grids[i].b, grids[i].width_y, grids[i].width_x, grids[i].steps, 0};
    }
    perks_batch_plan* p = new perks_batch_plan();
    p->dtype = dtype;
    p->opts = *o;
    PERKS_HIP_TRY(hipGetDevice(&p->device));
    rc = grid_sync_alloc(&p->gs);
    if (rc != PERKS_OK) {
        delete p;
        return rc;
    }
    int tile = 0;
    rc = resolve_tile(dtype, o, specs.data(), ngrids, &tile);
    if (rc != PERKS_OK) {
        (void)grid_sync_free(p->gs);
        delete p;
        return rc;
    }
    p->cfg = Config{tile, j2d_pass_levels(dtype, o->shape, o, tile), o->tune[3]};

    if (autotune && o->method != PERKS_METHOD_PLAIN) {
        // Scratch copies of (a prefix of) the batch's grids; contents do not matter for timing.
        const size_t esz = dtype == PERKS_F64 ? 8 : (dtype == PERKS_F32 ? 4 : 2);
        std::vector<J2dSpec> scratch;
        std::vector<void*> bufs;
        size_t used = 0;
        for (const J2dSpec& s : specs) {
            const size_t bytes = (size_t)s.ny * s.nx * esz;
            if (used + 2 * bytes > kTuneScratchBytes) break;
            void *a = nullptr, *b = nullptr;
            if (hipMalloc(&a, bytes) != hipSuccess) break;
            if (hipMalloc(&b, bytes) != hipSuccess) {
                (void)hipFree(a);
                break;
            }
            (void)hipMemsetAsync(a, 0, bytes, stream);
            (void)hipMemsetAsync(b, 0, bytes, stream);
            bufs.push_back(a);
            bufs.push_back(b);
            scratch.push_back(J2dSpec{a, b, s.ny, s.nx, s.steps, 0});
            used += 2 * bytes;
        }
        (void)hipStreamSynchronize(stream);
        double best = 0.0;
        Config best_c = p->cfg;
        const int occ_choices[] = {0, 1};
        const int level_choices[] = {1, 2, 3, 4, 6, 8, 12, 16};
        std::vector<int> tiles = {0, 1, kStreamTile};
        if (pair_applies(dtype, specs.data(), ngrids)) tiles.push_back(kPairTile);
        for (int tl : tiles)
            for (int L : level_choices)
                for (int occ : occ_choices) {
                    const Config c{tl, L, occ};
                    const double v = time_config(dtype, o, cf, c, scratch, sd->radius, kTuneSteps, stream, p->gs);
                    if (v > best) {
                        best = v;
                        best_c = c;
                    }
                }
        for (void* b : bufs) (void)hipFree(b);
        if (best > 0.0) {
            p->cfg = best_c;
            p->tuned = 1;
            p->tuned_gcells = best / 1e9;
        }
    }

    int tx, ty;
    tile_dims(dtype, p->cfg.tile, &tx, &ty);
    rc = build_work(specs.data(), ngrids, sd->radius, tx, ty, p->cfg.levels, &p->work);
    if (rc == PERKS_OK) rc = p->dw.upload(p->work, stream);
    if (rc == PERKS_OK) rc = prepare_launch(dtype, o, cf, p->cfg, p->work, p->dw, p->gs, &p->launch);
    if (rc == PERKS_OK && hipStreamSynchronize(stream) != hipSuccess) rc = PERKS_ERR_HIP;
    if (rc == PERKS_OK && !p->work.grids.empty()) {
        // Capture the launch(es) on a stream of our own (default priority).
        hipStream_t cap = nullptr;
        if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) rc = PERKS_ERR_HIP;
        if (rc == PERKS_OK) {
            if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = PERKS_ERR_HIP;
            const int erc = rc == PERKS_OK ? enqueue(p->launch, o, p->work, cap, false) : rc;
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(cap, &g);
            if (rc == PERKS_OK && (erc != PERKS_OK || ec != hipSuccess)) rc = erc != PERKS_OK ? erc : PERKS_ERR_HIP;
            if (rc == PERKS_OK) {
                p->graph = g;
                if (hipGraphInstantiate(&p->exec_plain, g, nullptr, nullptr, 0) != hipSuccess) rc = PERKS_ERR_HIP;
                if (rc == PERKS_OK && o->method != PERKS_METHOD_PLAIN) {
                    // the persistent grid: a cooperative replay, the ordinary one kept as the fallback
                    rc = mark_cooperative(g);
                    if (rc == PERKS_OK && hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0) != hipSuccess)
                        rc = PERKS_ERR_HIP;
                } else {
                    p->exec = p->exec_plain;
                    p->exec_plain = nullptr;
                }
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
            (void)hipStreamDestroy(cap);
        }
    }
    if (rc != PERKS_OK) {
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->exec_plain) (void)hipGraphExecDestroy(p->exec_plain);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        (void)p->dw.release();
        (void)grid_sync_free(p->gs);
        delete p;
        return rc;
    }
    *out = p;
    return PERKS_OK;
}

}  // namespace
}  // namespace perks

extern "C" {

int perks_jacobi2d_batch_plan(int dtype, const perks_grid2d* grids, int ngrids, const perks_stencil_opts* opts,
                              int autotune, void* stream, perks_batch_plan** plan)
{
    if (!grids || ngrids <= 0 || !opts || !plan) return PERKS_ERR_INVALID_ARG;
    if (dtype < PERKS_F64 || dtype > PERKS_BF16) return PERKS_ERR_INVALID_ARG;
    *plan = nullptr;
    return perks::plan_batch(dtype, grids, ngrids, opts, autotune, (hipStream_t)stream, plan);
}

int perks_jacobi2d_batch_run(perks_batch_plan* p, void* stream_v, perks_run_report* report)
{
    if (!p) return PERKS_ERR_INVALID_ARG;
    hipStream_t stream = (hipStream_t)stream_v;
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev != p->device) return PERKS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(p->run_mu);
    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    if (p->exec) {
        hipError_t e = hipGraphLaunch(p->exec, stream);
        if (e == hipErrorIllegalState && p->exec_plain) {
            // another process holds the device's cooperative queue: the
            // ordinary replay (co-residency then rests on an otherwise idle
            // device; the barrier's bounded wait reports a grid that is not)
            (void)hipGetLastError();
            e = hipGraphLaunch(p->exec_plain, stream);
        }
        PERKS_HIP_TRY(e);
    }
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (p->opts.method != PERKS_METHOD_PLAIN) {
        const int rc = perks::grid_sync_check(p->gs, stream);
        if (rc != PERKS_OK) return rc;
    }
    perks::fill_report(report, ms, p->work, p->launch, p->cfg, p->dtype);
    return PERKS_OK;
}

int perks_jacobi2d_batch_plan_info(const perks_batch_plan* p, int32_t* result_in_b, int ngrids, int32_t* config)
{
    if (!p) return PERKS_ERR_INVALID_ARG;
    if (result_in_b) {
        if (ngrids != (int)p->work.result_in_b.size()) return PERKS_ERR_INVALID_ARG;
        for (int i = 0; i < ngrids; ++i) result_in_b[i] = p->work.result_in_b[i];
    }
    if (config) {
        config[0] = p->cfg.tile;
        config[1] = p->cfg.levels;
        config[2] = p->launch.occ;
        config[3] = p->launch.grid;
        config[4] = p->tuned;
        config[5] = p->work.npass;
    }
    return PERKS_OK;
}

void perks_jacobi2d_batch_plan_destroy(perks_batch_plan* p)
{
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->exec_plain) (void)hipGraphExecDestroy(p->exec_plain);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    (void)p->dw.release();
    (void)perks::grid_sync_free(p->gs);
    delete p;
}

}  // extern "C"
