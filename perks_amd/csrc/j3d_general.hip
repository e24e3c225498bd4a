// j3d_general.hip -- host side of the shape-generic 3D kernel (j3d_general_kernel.h):
// tile and level planning, the per-step launches and the persistent launch.
//
// Replaces, for the 13-, 17-, 27-point and Poisson shapes, the reference's
// kernel3d_baseline / kernel3d_general paths (stencil/3dstencil/j3d-baseline.cu,
// j3d-general-kernels.cuh, j3d.cu:129-808); the 7-point shape has its own
// register-ring kernels (j3d_plain.hip, j3d_persistent_kernel.h).
//
// Two kernels serve these shapes: the register-streaming kernel
// (j3d_stream_kernel.h; one wave per 64 x 8 tile, planes in VGPRs; the
// default, 1.6-3.7x the other per step at 512^3) and the LDS-plane kernel
// (j3d_general_kernel.h; a workgroup per 64 x 16/32 tile, planes in LDS).
// opts->tune[4] picks one: 2 = streaming, 3 = LDS planes, 0 = default.
#include "j3d_general.h"

#include <algorithm>
#include <cstring>

#include "api_common.h"
#include "j3d_general_kernel.h"
#include "j3d_stream_kernel.h"

namespace perks {

namespace {

void* kernel_for(int dtype, int shape, int bc)
{
    switch (dtype) {
    case PERKS_F64: return j3g_kernel_f64(shape, bc);
    case PERKS_F32: return j3g_kernel_f32(shape, bc);
    case PERKS_F16: return j3g_kernel_f16(shape, bc);
    case PERKS_BF16: return j3g_kernel_bf16(shape, bc);
    default: return nullptr;
    }
}

void* stream_kernel_for(int dtype, int shape, int bc, int rows)
{
    switch (dtype) {
    case PERKS_F64: return j3s_kernel_f64(shape, bc, rows);
    case PERKS_F32: return j3s_kernel_f32(shape, bc, rows);
    case PERKS_F16: return j3s_kernel_f16(shape, bc, rows);
    case PERKS_BF16: return j3s_kernel_bf16(shape, bc, rows);
    default: return nullptr;
    }
}

// The default tile height of the streaming kernel per storage format, radius
// and method (see stream_rows).
int default_stream_rows(int dtype, int radius, int method)
{
    (void)dtype;
    (void)radius;
    (void)method;
    return j3s::kRows;
}

bool use_stream(const perks_stencil_opts* o, int radius)
{
    if (o->tune[4] == 2 || o->tune[4] == 4) return true;
    (void)radius;
    return o->tune[4] != 3;
}

// Tile rows per wave of the streaming kernel: tune[4] = 4 the tall tile
// (j3s::kRowsTall), 2 the short one (j3s::kRows); by default the tall tile
// where it measured faster.
int stream_rows(int dtype, int radius, const perks_stencil_opts* o)
{
    if (o->tune[4] == 4) return j3s::kRowsTall;
    if (o->tune[4] == 2) return j3s::kRows;
    return default_stream_rows(dtype, radius, o->method);
}

constexpr int kStreamItems = 16384;  // per-step launches: work items (waves) to aim for

int compute_bytes(int dtype) { return dtype == PERKS_F64 ? 8 : 4; }

void tile_of(int dtype, int* tx, int* ty)
{
    *tx = 64;
    *ty = dtype == PERKS_F64 ? 16 : 32;
}

// LDS the kernel needs for L levels: L rings of 2R+1 planes of the tile.
size_t lds_for(int dtype, int radius, int L)
{
    int tx, ty;
    tile_of(dtype, &tx, &ty);
    return (size_t)L * (2 * radius + 1) * tx * ty * compute_bytes(dtype);
}

constexpr size_t kLdsBudget = 64 * 1024;

// z-slab count minimising (rounds of items per workgroup) x (planes per item,
// the slab plus the 2RL ghost planes it reads beyond it).
int best_slabs(int nz, long long tiles_xy, int G, int ghost)
{
    int best = 1;
    double best_cost = 1e300;
    const int cap = std::min(nz, 512);
    for (int s = 1; s <= cap; ++s) {
        const long long items = tiles_xy * s;
        const long long rounds = (items + G - 1) / G;
        const double cost = (double)rounds * ((double)nz / s + ghost);
        if (cost < best_cost * 0.999) {
            best_cost = cost;
            best = s;
        }
    }
    return best;
}

}  // namespace

int j3g_levels(int dtype, int shape, const perks_stencil_opts* o)
{
    const ShapeDesc* sd = shape_by_id(shape);
    if (!sd) return 1;
    if (o->method == PERKS_METHOD_PLAIN) return 1;
    if (use_stream(o, sd->radius)) {
        const int lm = j3s::stream3_max_levels(sd->radius);
        return o->tune[2] > 0 ? std::min(o->tune[2], lm) : lm;
    }
    int tx, ty;
    tile_of(dtype, &tx, &ty);
    const int R = sd->radius;
    int L = o->tune[2] > 0 ? o->tune[2] : 4;
    while (L > 1 && lds_for(dtype, R, L) > kLdsBudget) --L;
    const int maxL = (std::min(tx, ty) - 2) / (2 * R);  // at least 2 valid cells per side
    return std::max(1, std::min(L, maxL));
}

int j3g_plan(int dtype, int nz, int ny, int nx, const perks_stencil_opts* o, const perks_device_caps* caps,
             int occupancy, perks_plan_info* out)
{
    const ShapeDesc* sd = shape_by_id(o->shape);
    if (!sd || sd->dims != 3 || !caps || !out) return PERKS_ERR_INVALID_ARG;
    int tx, ty;
    tile_of(dtype, &tx, &ty);
    const int R = sd->radius;
    const int L = j3g_levels(dtype, o->shape, o);
    if (use_stream(o, R)) {
        const int rows = stream_rows(dtype, R, o);
        const int VW = 64 - 2 * R * L, VH = rows - 2 * R * L;
        const long long tiles_xy = (long long)((nx + VW - 1) / VW) * ((ny + VH - 1) / VH);
        std::memset(out, 0, sizeof(*out));
        out->block = j3s::kWaves * 64;
        out->tile_x = 64;
        out->tile_y = rows;
        out->reserved[0] = L;
        int nzs = 1;
        long long groups = 0;
        if (o->method == PERKS_METHOD_PLAIN) {
            nzs = (int)std::max(1LL, std::min<long long>(nz, (kStreamItems + tiles_xy - 1) / tiles_xy));
            groups = (tiles_xy * nzs + j3s::kWaves - 1) / j3s::kWaves;
        } else {
            if (occupancy <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
            int occ = std::min(occupancy, kCoopMaxWavesPerCu / j3s::kWaves);
            if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
            if (o->tune[3] > 0 && o->tune[3] < occ) occ = o->tune[3];
            const int G = caps->cu_count * occ;
            nzs = best_slabs(nz, tiles_xy, G * j3s::kWaves, 2 * R * L);
            groups = std::min<long long>(G, (tiles_xy * nzs + j3s::kWaves - 1) / j3s::kWaves);
            out->occupancy = occ;
        }
        if (groups > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
        out->grid[0] = (int)groups;
        out->grid[1] = out->grid[2] = 1;
        out->reserved[1] = nzs;
        return PERKS_OK;
    }
    const int VW = tx - 2 * R * L, VH = ty - 2 * R * L;
    const long long tiles_xy = (long long)((nx + VW - 1) / VW) * ((ny + VH - 1) / VH);
    std::memset(out, 0, sizeof(*out));
    out->block = j3g::kThreads;
    out->tile_x = tx;
    out->tile_y = ty;
    out->lds_bytes = (int)lds_for(dtype, R, L);
    out->reserved[0] = L;
    if (o->method == PERKS_METHOD_PLAIN) {
        const int nzs = (int)std::max(1LL, std::min<long long>(nz, (8LL * caps->cu_count + tiles_xy - 1) / tiles_xy));
        out->grid[0] = (int)std::min<long long>(tiles_xy * nzs, 0x7fffffffLL);
        out->reserved[1] = nzs;
    } else {
        if (occupancy <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
        int occ = std::min(occupancy, std::max(1, kCoopMaxWavesPerCu / (j3g::kThreads / 64)));
        if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
        if (o->tune[3] > 0 && o->tune[3] < occ) occ = o->tune[3];
        const int G = caps->cu_count * occ;
        const int nzs = best_slabs(nz, tiles_xy, G, 2 * R * L);
        out->occupancy = occ;
        out->grid[0] = (int)std::min<long long>(G, tiles_xy * nzs);
        out->reserved[1] = nzs;
    }
    out->grid[1] = out->grid[2] = 1;
    return PERKS_OK;
}

void* j3g_fn(int dtype, const perks_stencil_opts* o, int* block, size_t* lds)
{
    const ShapeDesc* sd = shape_by_id(o->shape);
    if (!sd || sd->dims != 3) return nullptr;
    const bool streaming = use_stream(o, sd->radius);
    if (block) *block = streaming ? j3s::kWaves * 64 : j3g::kThreads;
    if (lds) *lds = streaming ? 0 : lds_for(dtype, sd->radius, j3g_levels(dtype, o->shape, o));
    return streaming ? stream_kernel_for(dtype, o->shape, o->boundary, stream_rows(dtype, sd->radius, o))
                     : kernel_for(dtype, o->shape, o->boundary);
}

int j3g_run(int dtype, void* d_a, void* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
            const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep)
{
    const ShapeDesc* sd = shape_by_id(o->shape);
    if (!sd || sd->dims != 3) return PERKS_ERR_INVALID_ARG;
    const int R = sd->radius;
    const bool streaming = use_stream(o, R);
    const int rows = streaming ? stream_rows(dtype, R, o) : 0;
    void* fn = streaming ? stream_kernel_for(dtype, o->shape, o->boundary, rows) : kernel_for(dtype, o->shape, o->boundary);
    if (!fn) return PERKS_ERR_UNSUPPORTED;
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    int tx, ty;
    tile_of(dtype, &tx, &ty);
    if (streaming) {
        tx = 64;
        ty = rows;
    }
    const int L = j3g_levels(dtype, o->shape, o);
    const size_t lds = streaming ? 0 : lds_for(dtype, R, L);
    const int threads = streaming ? j3s::kWaves * 64 : j3g::kThreads;
    const int per_group = streaming ? j3s::kWaves : 1;  // work items per workgroup
    const int VW = tx - 2 * R * L, VH = ty - 2 * R * L;
    const long long tiles_xy = (long long)((nx + VW - 1) / VW) * ((ny + VH - 1) / VH);

    j3g::J3gArgs a;
    std::memset(&a, 0, sizeof(a));
    a.a = d_a;
    a.b = d_b;
    a.nz = nz;
    a.ny = ny;
    a.nx = nx;
    a.steps = steps;
    a.levels = L;
    a.has_div = cf.divisor != 0.0;
    a.sr = (dtype == PERKS_F16 || dtype == PERKS_BF16) && o->rounding == PERKS_ROUND_STOCHASTIC;
    a.seed = o->round_seed;
    KCoefs kc = cf;
    SyncLease lease;
    rc = lease.acquire();
    if (rc != PERKS_OK) return rc;
    GridSync* gs = lease.gs;
    a.sync = gs;
    a.timeout = grid_sync_timeout();

    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    int grid = 0, occ = 0;
    const int npass = steps > 0 ? (steps + L - 1) / L : 0;
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    if (o->method == PERKS_METHOD_PLAIN) {
        // One launch per step.  LDS kernel: every (tile, slab) item its own
        // workgroup, enough slabs for about eight workgroups per CU; streaming
        // kernel: one item per wave, enough slabs for kStreamItems items.
        const long long want = streaming ? kStreamItems : 8LL * caps.cu_count;
        a.nzs = (int)std::max(1LL, std::min<long long>(nz, (want + tiles_xy - 1) / tiles_xy));
        const long long nitems = tiles_xy * a.nzs;
        if (nitems > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
        grid = (int)((nitems + per_group - 1) / per_group);
        a.npass = 1;
        for (int s = 0; s < steps; ++s) {
            a.step_offset = s;
            void* args[] = {&a, &kc};
            PERKS_HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, stream));
        }
    } else {
        PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, threads, lds));
        if (occ <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
        occ = std::min(occ, std::max(1, kCoopMaxWavesPerCu / (threads / 64)));
        if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
        if (o->tune[3] > 0 && o->tune[3] < occ) occ = o->tune[3];
        const int G = caps.cu_count * occ;
        a.nzs = best_slabs(nz, tiles_xy, G * per_group, 2 * R * L);
        grid = (int)std::min<long long>(G, (tiles_xy * a.nzs + per_group - 1) / per_group);
        a.npass = npass;
        a.step_offset = 0;
        void* args[] = {&a, &kc};
        if (steps > 0)
            PERKS_HIP_TRY(launch_persistent(fn, dim3(grid), dim3(threads), args, lds, stream));
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
        rep->gcells_per_s = ms > 0 ? (double)steps * nz * ny * (double)nx / ms / 1e6 : 0.0;
        rep->grid[0] = grid;
        rep->grid[1] = rep->grid[2] = 1;
        rep->block = threads;
        rep->lds_bytes = (int)lds;
        rep->lds_planes = streaming ? 0 : L * (2 * R + 1);
        rep->reg_planes = streaming ? (L + 1) * (2 * R + 1) : 0;
        rep->occupancy = occ;
        rep->result_in_b = (o->method == PERKS_METHOD_PLAIN ? steps : npass) % 2;
        rep->reserved[0] = L;
        rep->reserved[1] = npass;
    }
    return PERKS_OK;
}

}  // namespace perks
