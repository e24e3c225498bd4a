// j3d_tune.hip -- configuration table and autotuner of the 3D path.
//
// The reference fixes its register-cache depth per configuration at build
// time: ProcessCompilerLog.py (lines 47-75) reads nvcc's register and spill
// report for every template instantiation, and ProcessRst2cuh.py (lines 5-51)
// writes the deepest non-spilling depth into the regfolder<> table of
// stencil/3dstencil/perksconfig.cuh, which j3d_iterative then reads.
//
// Here every configuration (per-step kernel, each persistent tile variant at
// each pass depth, register-streaming or LDS-plane kernel for the other
// shapes) is already compiled into the library.  perks_j3d_kernel_table reads
// the same figures from the loaded code objects -- VGPRs, scratch (spill)
// bytes, LDS, the occupancy the launch would use -- and marks what cannot run
// on the problem; perks_j3d_autotune times every runnable configuration on
// the caller's field and returns the fastest in the options.  Every
// configuration computes the same bits, so tuning changes only the speed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "api_common.h"
#include "j3d_general.h"
#include "j3d_kernels.h"
#include "j3d_persistent.h"
#include "j3d_resident.h"
#include "planner.h"

namespace perks {
namespace {

constexpr int kDefaultSampleSteps = 12;  // a multiple of every pass depth (1..4)

struct Candidate {
    int method;
    int variant;
    int levels;
};

int elem_size(int dtype) { return dtype == PERKS_F64 ? 8 : (dtype == PERKS_F32 ? 4 : 2); }

// The 7-point star runs on its own kernels unless tune[5] = 1 sends it to
// the shape-generic ones (the same rule as the run entry points).
bool seven_point(const perks_stencil_opts* o) { return o->shape == PERKS_SHAPE_3D_STAR_R1 && o->tune[5] != 1; }

perks_stencil_opts with(const perks_stencil_opts* base, const Candidate& c)
{
    perks_stencil_opts o = *base;
    o.method = c.method;
    o.tune[2] = c.method == PERKS_METHOD_PLAIN ? 0 : c.levels;
    o.tune[4] = c.variant;
    return o;
}

std::vector<Candidate> candidates(int dtype, const perks_stencil_opts* o)
{
    std::vector<Candidate> c;
    if (seven_point(o)) {
        c.push_back({PERKS_METHOD_PLAIN, 0, 1});
        const int nv = j3d_persistent_variants();
        for (int v = 0; v < nv; ++v) {
            for (int l = 1; l <= 4; ++l) c.push_back({PERKS_METHOD_PERSISTENT, v, l});
        }
        c.push_back({PERKS_METHOD_PERSISTENT, kResidentVariant, 1});  // resident tiles: one step per barrier
        return c;
    }
    for (int k : {2, 4, 3}) {  // register streaming (short and tall tiles), LDS planes
        c.push_back({PERKS_METHOD_PLAIN, k, 1});
        perks_stencil_opts p = with(o, {PERKS_METHOD_PERSISTENT, k, 0});
        const int maxl = j3g_levels(dtype, o->shape, &p);
        for (int l = 1; l <= maxl; ++l) c.push_back({PERKS_METHOD_PERSISTENT, k, l});
    }
    return c;
}

int describe(int dtype, int nz, int ny, int nx, const perks_stencil_opts* base, const Candidate& c,
             const perks_device_caps& caps, const KCoefs& cf, perks_kernel_config* out)
{
    const bool has_div = cf.divisor != 0.0;
    std::memset(out, 0, sizeof(*out));
    out->method = c.method;
    out->variant = c.variant;
    out->levels = c.levels;
    const perks_stencil_opts o = with(base, c);
    const bool sr = (dtype == PERKS_F16 || dtype == PERKS_BF16) && o.rounding == PERKS_ROUND_STOCHASTIC;
    void* fn = nullptr;
    int block = 256;
    size_t lds = 0;
    if (seven_point(&o)) {
        if (c.method == PERKS_METHOD_PLAIN) {
            if (o.tune[6] == 1) return out->status = PERKS_ERR_UNSUPPORTED;  // contracted: persistent only
            fn = elem_size(dtype) == 2 ? j3d_star1_plain16_fn(dtype, o.boundary, has_div, sr)
                                       : j3d_star1_plain_fn(dtype, o.boundary, has_div, o.tune[1] != 2);
            out->tile_x = elem_size(dtype) == 2 ? 128 : 64;
            out->tile_y = 32;
        } else if (c.variant == kResidentVariant) {
            if (o.tune[6] == 1) return out->status = PERKS_ERR_UNSUPPORTED;  // contracted: ring kernel only
            fn = j3d_resident_fn(dtype, o.boundary, has_div, &block);
        } else {
            fn = j3d_persistent_fn(dtype, o.boundary, has_div, c.variant, sr, j3d_persistent_update(dtype, &o, cf),
                                   &block);
        }
    } else {
        fn = j3g_fn(dtype, &o, &block, &lds);
    }
    out->block = block;
    if (!fn) return out->status = PERKS_ERR_UNSUPPORTED;
    hipFuncAttributes at;
    PERKS_HIP_TRY(hipFuncGetAttributes(&at, fn));
    out->vgprs = at.numRegs;
    out->spill_bytes = (int)at.localSizeBytes;
    out->lds_bytes = (int)(at.sharedSizeBytes + lds);
    int occ = 0;
    PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, block, lds));
    out->occupancy = occ;
    if (at.localSizeBytes > 0) return out->status = PERKS_ERR_SPILL;
    if (occ <= 0) return out->status = PERKS_ERR_NOT_LAUNCHABLE;
    perks_plan_info pl;
    int rc = PERKS_OK;
    if (seven_point(&o)) {
        if (c.method == PERKS_METHOD_PLAIN) {
            if (elem_size(dtype) == 2 && (nx & 1)) rc = PERKS_ERR_UNSUPPORTED_SIZE;
        } else {
            rc = plan_persistent(dtype, nz, ny, nx, &o, &caps, occ, &pl);
            if (rc == PERKS_OK) {
                out->occupancy = pl.occupancy;
                out->tile_x = pl.tile_x;
                out->tile_y = pl.tile_y;
            }
        }
    } else {
        rc = j3g_plan(dtype, nz, ny, nx, &o, &caps, occ, &pl);
        if (rc == PERKS_OK) {
            if (c.method == PERKS_METHOD_PERSISTENT) out->occupancy = pl.occupancy;
            out->tile_x = pl.tile_x;
            out->tile_y = pl.tile_y;
        }
    }
    if (rc == PERKS_OK && c.method == PERKS_METHOD_PERSISTENT && !caps.cooperative) rc = PERKS_ERR_NOT_LAUNCHABLE;
    return out->status = rc;
}

int check_problem(int dtype, int nz, int ny, int nx, const perks_stencil_opts* o)
{
    if (!o) return PERKS_ERR_INVALID_ARG;
    if (dtype < PERKS_F64 || dtype > PERKS_BF16) return PERKS_ERR_INVALID_ARG;
    if (nz <= 0 || ny <= 0 || nx <= 0) return PERKS_ERR_INVALID_ARG;
    const ShapeDesc* s = shape_by_id(o->shape);
    if (!s || s->dims != 3) return PERKS_ERR_INVALID_ARG;
    if (o->boundary != PERKS_BC_CLAMP && o->boundary != PERKS_BC_ZERO) return PERKS_ERR_INVALID_ARG;
    return PERKS_OK;
}

int build_table(int dtype, int nz, int ny, int nx, const perks_stencil_opts* o, std::vector<Candidate>* cands,
                std::vector<perks_kernel_config>* rows)
{
    int rc = check_problem(dtype, nz, ny, nx, o);
    if (rc != PERKS_OK) return rc;
    KCoefs cf;
    rc = resolve_coefs(o, &cf, nullptr);
    if (rc != PERKS_OK) return rc;
    perks_device_caps caps;
    rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    *cands = candidates(dtype, o);
    rows->resize(cands->size());
    for (size_t i = 0; i < cands->size(); ++i) {
        rc = describe(dtype, nz, ny, nx, o, (*cands)[i], caps, cf, &(*rows)[i]);
        if (rc == PERKS_ERR_HIP) return rc;
    }
    return PERKS_OK;
}

void copy_out(const std::vector<perks_kernel_config>& rows, perks_kernel_config* table, int capacity, int* count)
{
    if (count) *count = (int)rows.size();
    if (!table) return;
    const int n = std::min<int>(capacity, (int)rows.size());
    for (int i = 0; i < n; ++i) table[i] = rows[i];
}

}  // namespace
}  // namespace perks

extern "C" {

int perks_j3d_kernel_table(int dtype, int height, int width_y, int width_x, const perks_stencil_opts* opts,
                           perks_kernel_config* table, int capacity, int* count)
{
    if (capacity < 0) return PERKS_ERR_INVALID_ARG;
    std::vector<perks::Candidate> cands;
    std::vector<perks_kernel_config> rows;
    const int rc = perks::build_table(dtype, height, width_y, width_x, opts, &cands, &rows);
    if (rc != PERKS_OK) return rc;
    perks::copy_out(rows, table, capacity, count);
    return PERKS_OK;
}

int perks_j3d_autotune(int dtype, const void* d_in, int height, int width_y, int width_x, int sample_steps,
                       perks_stencil_opts* opts, perks_kernel_config* table, int capacity, int* count,
                       void* stream)
{
    if (!d_in || capacity < 0) return PERKS_ERR_INVALID_ARG;
    std::vector<perks::Candidate> cands;
    std::vector<perks_kernel_config> rows;
    int rc = perks::build_table(dtype, height, width_y, width_x, opts, &cands, &rows);
    if (rc != PERKS_OK) return rc;
    const int steps = sample_steps > 0 ? sample_steps : perks::kDefaultSampleSteps;
    const size_t bytes = (size_t)height * width_y * (size_t)width_x * perks::elem_size(dtype);
    hipStream_t s = (hipStream_t)stream;
    void* a = nullptr;
    void* b = nullptr;
    PERKS_HIP_TRY(hipMalloc(&a, bytes));
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return PERKS_ERR_HIP;
    }
    int best = -1;
    for (size_t i = 0; i < cands.size() && rc == PERKS_OK; ++i) {
        if (rows[i].status != PERKS_OK) continue;
        const perks_stencil_opts o = perks::with(opts, cands[i]);
        perks_run_report rep;
        // One untimed pass first (first-launch costs), then the timed run,
        // each from the caller's field.
        const int warm = cands[i].method == PERKS_METHOD_PLAIN ? 1 : cands[i].levels;
        for (int pass = 0; pass < 2; ++pass) {
            if (hipMemcpyAsync(a, d_in, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) {
                rc = PERKS_ERR_HIP;
                break;
            }
            const int r = perks_j3d_device(dtype, a, b, height, width_y, width_x, pass == 0 ? warm : steps, &o,
                                           stream, &rep);
            if (r == PERKS_ERR_HIP || r == PERKS_ERR_BARRIER_TIMEOUT) {
                rc = r;
                break;
            }
            if (r != PERKS_OK) {
                rows[i].status = r;
                break;
            }
        }
        if (rc != PERKS_OK || rows[i].status != PERKS_OK) continue;
        rows[i].gcells_per_s = rep.gcells_per_s;
        if (best < 0 || rep.gcells_per_s > rows[best].gcells_per_s) best = (int)i;
    }
    (void)hipFree(a);
    (void)hipFree(b);
    if (rc != PERKS_OK) return rc;
    perks::copy_out(rows, table, capacity, count);
    if (best < 0) return PERKS_ERR_UNSUPPORTED;
    opts->method = cands[best].method;
    opts->tune[2] = cands[best].method == PERKS_METHOD_PLAIN ? 0 : cands[best].levels;
    opts->tune[4] = cands[best].variant;
    return PERKS_OK;
}

}  // extern "C"
