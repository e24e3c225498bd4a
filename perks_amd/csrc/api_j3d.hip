// api_j3d.hip -- C ABI of the library core and of the 3D Jacobi path.
//
// perks_j3d_iterative_* reproduce the contract of the reference's
// j3d_iterative<REAL> (stencil/3dstencil/j3d.cu:129-808): host buffers in and
// out, device memory allocated and freed by the callee, the result taken from
// the buffer the step parity selects (j3d.cu:786-798), the timed region being
// the launch(es) only (j3d.cu:720-750), and the planner's refusals returned as
// the same negative codes.  perks_j3d_device_* run the same work on buffers
// already resident in HBM.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "api_common.h"
#include "j3d_kernels.h"
#include "storage16.h"
#include "j3d_general.h"
#include "j3d_persistent.h"
#include "planner.h"
#include "j2d.h"
#include "cg.h"

namespace perks {

int device_caps_cached(perks_device_caps* caps)
{
    static std::mutex mu;
    static int cached_dev = -1;
    static perks_device_caps cached;
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (cached_dev != dev) {
        int rc = perks_query_device(dev, &cached);
        if (rc != PERKS_OK) return rc;
        cached_dev = dev;
    }
    *caps = cached;
    return PERKS_OK;
}

namespace {

template <typename T>
int run_plain(T* d_a, T* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf, const perks_stencil_opts* o,
              hipStream_t stream, perks_run_report* rep)
{
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    perks_plan_info pl;
    rc = plan_plain(nz, ny, nx, caps.cu_count, &pl);
    if (rc != PERKS_OK) return rc;
    const int zchunk = o->tune[0] > 0 ? o->tune[0] : pl.reserved[0];
    const int variant = o->tune[1] == 2 ? 0 : 1;  // nontemporal stores unless asked otherwise
    const int bc = o->boundary;
    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PlainGeom geom{0, 0, 0};
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    T* src = d_a;
    T* dst = d_b;
    for (int s = 0; s < steps && rc == PERKS_OK; ++s) {
        rc = launch_j3d_star1_plain<T>(src, dst, nz, ny, nx, cf, bc, zchunk, stream, &geom, variant);
        T* t = src;
        src = dst;
        dst = t;
    }
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (rc != PERKS_OK) return rc;
    if (rep) {
        rep->elapsed_ms = ms;
        rep->gcells_per_s = ms > 0 ? (double)steps * nz * ny * (double)nx / ms / 1e6 : 0.0;
        rep->grid[0] = geom.grid;
        rep->grid[1] = 1;
        rep->grid[2] = 1;
        rep->block = geom.block;
        rep->lds_bytes = 0;
        rep->lds_planes = 0;
        rep->reg_planes = 0;
        rep->occupancy = 0;
        rep->result_in_b = steps % 2 == 1;
    }
    return PERKS_OK;
}

// Every shape but the 7-point star runs on the shape-generic kernel; tune[5] = 1
// sends the 7-point star there too.
bool uses_general(const perks_stencil_opts* o)
{
    return o->shape != PERKS_SHAPE_3D_STAR_R1 || o->tune[5] == 1;
}

template <typename T>
int validate(const T* a, const T* b, int nz, int ny, int nx, int steps, const perks_stencil_opts* o)
{
    if (!a || !b || !o) return PERKS_ERR_INVALID_ARG;
    if (nz <= 0 || ny <= 0 || nx <= 0 || steps < 0) return PERKS_ERR_INVALID_ARG;
    // One z-plane is addressed through a buffer descriptor with 32-bit byte
    // offsets, and offset 2^31 marks "no access": a plane must stay below 2 GiB.
    if ((long long)ny * nx * (long long)sizeof(T) >= (long long)kOob) return PERKS_ERR_UNSUPPORTED_SIZE;
    if (o->boundary != PERKS_BC_CLAMP && o->boundary != PERKS_BC_ZERO) return PERKS_ERR_INVALID_ARG;
    if (o->method != PERKS_METHOD_PLAIN && o->method != PERKS_METHOD_PERSISTENT) return PERKS_ERR_INVALID_ARG;
    if (o->rounding != PERKS_ROUND_NEAREST && o->rounding != PERKS_ROUND_STOCHASTIC) return PERKS_ERR_INVALID_ARG;
    const ShapeDesc* s = shape_by_id(o->shape);
    if (!s || s->dims != 3) return PERKS_ERR_INVALID_ARG;
    // tune[6] = 1: the contracted 7-point update, persistent kernel, FP64/FP32 storage
    if (o->tune[6] != 0) {
        if (o->tune[6] != 1) return PERKS_ERR_INVALID_ARG;
        if (sizeof(T) == 2 || uses_general(o) || o->method != PERKS_METHOD_PERSISTENT) return PERKS_ERR_UNSUPPORTED;
    }
    if (o->tune[7] < 0) return PERKS_ERR_INVALID_ARG;  // z-slabs per tile (0: the planner's)
    return PERKS_OK;
}

template <typename T>
int j3d_device(T* d_a, T* d_b, int nz, int ny, int nx, int steps, const perks_stencil_opts* o, void* stream,
               perks_run_report* rep)
{
    int rc = validate<T>(d_a, d_b, nz, ny, nx, steps, o);
    if (rc != PERKS_OK) return rc;
    KCoefs cf;
    rc = resolve_coefs(o, &cf, nullptr);
    if (rc != PERKS_OK) return rc;
    if (rep) std::memset(rep, 0, sizeof(*rep));
    hipStream_t s = (hipStream_t)stream;
    if (uses_general(o))
        return j3g_run(sizeof(T) == 8 ? PERKS_F64 : PERKS_F32, d_a, d_b, nz, ny, nx, steps, cf, o, s, rep);
    if (o->method == PERKS_METHOD_PLAIN) return run_plain<T>(d_a, d_b, nz, ny, nx, steps, cf, o, s, rep);
    return run_persistent<T>(d_a, d_b, nz, ny, nx, steps, cf, o, s, rep);
}

// FP16/BF16 storage (16-bit patterns), FP32 arithmetic.
int run_plain16(int fmt, uint16_t* d_a, uint16_t* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
                const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep)
{
    if ((nx & 1) != 0) return PERKS_ERR_UNSUPPORTED_SIZE;
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    perks_plan_info pl;
    rc = plan_plain(nz, ny, nx / 2, caps.cu_count, &pl);  // a lane holds two cells
    if (rc != PERKS_OK) return rc;
    const int zchunk = o->tune[0] > 0 ? o->tune[0] : pl.reserved[0];
    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PlainGeom geom{0, 0, 0};
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    uint16_t* src = d_a;
    uint16_t* dst = d_b;
    for (int s = 0; s < steps && rc == PERKS_OK; ++s) {
        const bool sr = o->rounding == PERKS_ROUND_STOCHASTIC;
        rc = launch_j3d_star1_plain16(fmt, src, dst, nz, ny, nx, cf, o->boundary, zchunk, sr,
                                      sr_key((uint32_t)s, o->round_seed), stream, &geom);
        uint16_t* t = src;
        src = dst;
        dst = t;
    }
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (rc != PERKS_OK) return rc;
    if (rep) {
        rep->elapsed_ms = ms;
        rep->gcells_per_s = ms > 0 ? (double)steps * nz * ny * (double)nx / ms / 1e6 : 0.0;
        rep->grid[0] = geom.grid;
        rep->grid[1] = rep->grid[2] = 1;
        rep->block = geom.block;
        rep->result_in_b = steps % 2 == 1;
    }
    return PERKS_OK;
}

int j3d_device16(int fmt, uint16_t* d_a, uint16_t* d_b, int nz, int ny, int nx, int steps,
                 const perks_stencil_opts* o, void* stream, perks_run_report* rep)
{
    int rc = validate<uint16_t>(d_a, d_b, nz, ny, nx, steps, o);
    if (rc != PERKS_OK) return rc;
    KCoefs cf;
    rc = resolve_coefs(o, &cf, nullptr);
    if (rc != PERKS_OK) return rc;
    if (rep) std::memset(rep, 0, sizeof(*rep));
    hipStream_t s = (hipStream_t)stream;
    if (uses_general(o)) return j3g_run(fmt, d_a, d_b, nz, ny, nx, steps, cf, o, s, rep);
    if (o->method == PERKS_METHOD_PLAIN) return run_plain16(fmt, d_a, d_b, nz, ny, nx, steps, cf, o, s, rep);
    return run_persistent16(fmt, d_a, d_b, nz, ny, nx, steps, cf, o, s, rep);
}

template <typename T>
int j3d_host(const T* h_in, int nz, int ny, int nx, T* h_out, int steps, const perks_stencil_opts* o,
             perks_run_report* rep)
{
    int rc = validate<T>(h_in, h_out, nz, ny, nx, steps, o);
    if (rc != PERKS_OK) return rc;
    const size_t n = (size_t)nz * ny * nx;
    const size_t bytes = n * sizeof(T);
    T* d_a = nullptr;
    T* d_b = nullptr;
    PERKS_HIP_TRY(hipMalloc(&d_a, bytes));
    if (hipMalloc(&d_b, bytes) != hipSuccess) {
        (void)hipFree(d_a);
        return PERKS_ERR_HIP;
    }
    perks_run_report local;
    perks_run_report* r = rep ? rep : &local;
    do {
        if (o->warmup) {
            // Untimed launches on the scratch buffers before the input is
            // uploaded (the reference warms up on its output arrays, j3d.cu:655-718).
            const int ws = o->warmup_steps > 0 ? o->warmup_steps : 10;
            if (hipMemset(d_a, 0, bytes) != hipSuccess) { rc = PERKS_ERR_HIP; break; }
            rc = j3d_device<T>(d_a, d_b, nz, ny, nx, ws, o, nullptr, r);
            if (rc != PERKS_OK) break;
        }
        if (hipMemcpy(d_a, h_in, bytes, hipMemcpyHostToDevice) != hipSuccess) { rc = PERKS_ERR_HIP; break; }
        rc = j3d_device<T>(d_a, d_b, nz, ny, nx, steps, o, nullptr, r);
        if (rc != PERKS_OK) break;
        const T* res = r->result_in_b ? d_b : d_a;
        if (hipMemcpy(h_out, res, bytes, hipMemcpyDeviceToHost) != hipSuccess) { rc = PERKS_ERR_HIP; break; }
    } while (0);
    (void)hipFree(d_a);
    (void)hipFree(d_b);
    return rc;
}

}  // namespace
}  // namespace perks

extern "C" {

const char* perks_version(void) { return "perks_amd 0.1 (gfx950)"; }

const char* perks_error_string(int code)
{
    switch (code) {
    case PERKS_OK: return "ok";
    case PERKS_ERR_NO_LDS_FOR_CACHE: return "no LDS space left for cached planes";
    case PERKS_ERR_UNSUPPORTED_SIZE: return "domain size not supported";
    case PERKS_ERR_NOT_LAUNCHABLE: return "kernel is not launchable (occupancy 0)";
    case PERKS_ERR_HEIGHT_TOO_SMALL: return "height too small for the persistent grid";
    case PERKS_ERR_SPILL: return "cache depth would spill registers";
    case PERKS_ERR_INVALID_ARG: return "invalid argument";
    case PERKS_ERR_HIP: return "HIP runtime error";
    case PERKS_ERR_NO_DEVICE: return "no HIP device";
    case PERKS_ERR_UNSUPPORTED: return "shape/dtype/method combination not built";
    case PERKS_ERR_NOT_CONVERGED: return "solver did not converge";
    case PERKS_ERR_BARRIER_TIMEOUT: return "grid barrier timed out (launch abandoned)";
    default: return "unknown error";
    }
}

int perks_query_device(int device, perks_device_caps* caps)
{
    if (!caps) return PERKS_ERR_INVALID_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PERKS_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return PERKS_ERR_INVALID_ARG;
    hipDeviceProp_t p;
    PERKS_HIP_TRY(hipGetDeviceProperties(&p, device));
    std::memset(caps, 0, sizeof(*caps));
    caps->cu_count = p.multiProcessorCount;
    caps->lds_per_cu = (int)p.maxSharedMemoryPerMultiProcessor;
    caps->lds_per_block = (int)p.sharedMemPerBlock;
    caps->vgprs_per_simd = 512;
    caps->simds_per_cu = 4;
    caps->warp_size = p.warpSize;
    caps->cooperative = p.cooperativeLaunch;
    return PERKS_OK;
}

void perks_stencil_opts_init(perks_stencil_opts* o, int shape)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->shape = shape;
    o->boundary = PERKS_BC_CLAMP;
    o->method = PERKS_METHOD_PERSISTENT;
    o->use_ref_coefs = 1;
    o->reg_cache_planes = -1;
    const perks::ShapeDesc* s = perks::shape_by_id(shape);
    if (s) {
        for (int g = 0; g < s->ngroups; ++g) o->coef[g] = s->g[g].coef;
        o->divisor = s->divisor;
    }
}

int perks_shape_info(int shape, int32_t* dims, int32_t* radius, int32_t* ngroups, double* coef, double* divisor,
                     int32_t* npts, int32_t* prod_f64, int8_t* offsets)
{
    const perks::ShapeDesc* s = perks::shape_by_id(shape);
    if (!s) return PERKS_ERR_INVALID_ARG;
    if (dims) *dims = s->dims;
    if (radius) *radius = s->radius;
    if (ngroups) *ngroups = s->ngroups;
    if (divisor) *divisor = s->divisor;
    for (int g = 0; g < s->ngroups; ++g) {
        if (coef) coef[g] = s->g[g].coef;
        if (npts) npts[g] = s->g[g].npts;
        if (prod_f64) prod_f64[g] = s->g[g].prod_f64;
        if (offsets)
            for (int p = 0; p < perks::kMaxPts; ++p)
                for (int a = 0; a < 3; ++a) offsets[(g * perks::kMaxPts + p) * 3 + a] = s->g[g].off[p][a];
    }
    return PERKS_OK;
}

int perks_release(void)
{
    const int a = perks::release_persistent_scratch();
    const int b = perks::release_j2d_scratch();
    const int c = perks::release_cg_scratch();
    return a != PERKS_OK ? a : (b != PERKS_OK ? b : c);
}

int perks_j3d_iterative_f64(const double* h_in, int height, int width_y, int width_x, double* h_out, int steps,
                            const perks_stencil_opts* opts, perks_run_report* report)
{
    return perks::j3d_host<double>(h_in, height, width_y, width_x, h_out, steps, opts, report);
}

int perks_j3d_iterative_f32(const float* h_in, int height, int width_y, int width_x, float* h_out, int steps,
                            const perks_stencil_opts* opts, perks_run_report* report)
{
    return perks::j3d_host<float>(h_in, height, width_y, width_x, h_out, steps, opts, report);
}

int perks_j3d_device_f64(double* d_a, double* d_b, int height, int width_y, int width_x, int steps,
                         const perks_stencil_opts* opts, void* stream, perks_run_report* report)
{
    return perks::j3d_device<double>(d_a, d_b, height, width_y, width_x, steps, opts, stream, report);
}

int perks_j3d_device_f32(float* d_a, float* d_b, int height, int width_y, int width_x, int steps,
                         const perks_stencil_opts* opts, void* stream, perks_run_report* report)
{
    return perks::j3d_device<float>(d_a, d_b, height, width_y, width_x, steps, opts, stream, report);
}

int perks_j3d_device(int dtype, void* d_a, void* d_b, int height, int width_y, int width_x, int steps,
                     const perks_stencil_opts* opts, void* stream, perks_run_report* report)
{
    switch (dtype) {
    case PERKS_F64:
        return perks::j3d_device<double>((double*)d_a, (double*)d_b, height, width_y, width_x, steps, opts, stream,
                                         report);
    case PERKS_F32:
        return perks::j3d_device<float>((float*)d_a, (float*)d_b, height, width_y, width_x, steps, opts, stream,
                                        report);
    case PERKS_F16:
    case PERKS_BF16:
        return perks::j3d_device16(dtype, (uint16_t*)d_a, (uint16_t*)d_b, height, width_y, width_x, steps, opts,
                                   stream, report);
    default: return PERKS_ERR_INVALID_ARG;
    }
}

}  // extern "C"
