// api_j2d.hip -- C ABI of the 2D Jacobi path.
//
// perks_jacobi2d_iterative_* keep the contract of the reference's
// jacobi_iterative<REAL> (stencil/2dstencil/jacobi-2d.cu:369-865): host
// buffers in and out, device memory owned by the callee, the result taken
// from the buffer the pass parity selects, timing of the launches only.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "api_common.h"
#include "j2d.h"

namespace perks {
namespace {

int validate2(const perks_stencil_opts* o, int wy, int wx, int steps)
{
    if (!o || wy <= 0 || wx <= 0 || steps < 0) return PERKS_ERR_INVALID_ARG;
    if ((long long)wy * wx > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
    const ShapeDesc* s = shape_by_id(o->shape);
    if (!s || s->dims != 2) return PERKS_ERR_INVALID_ARG;
    if (o->method != PERKS_METHOD_PLAIN && o->method != PERKS_METHOD_PERSISTENT) return PERKS_ERR_INVALID_ARG;
    return PERKS_OK;
}

size_t esize(int dtype) { return dtype == PERKS_F64 ? 8 : dtype == PERKS_F32 ? 4 : 2; }

int device2(int dtype, void* a, void* b, int wy, int wx, int steps, const perks_stencil_opts* o, void* stream,
            perks_run_report* rep)
{
    int rc = validate2(o, wy, wx, steps);
    if (rc != PERKS_OK) return rc;
    if (!a || !b) return PERKS_ERR_INVALID_ARG;
    if (dtype < PERKS_F64 || dtype > PERKS_BF16) return PERKS_ERR_INVALID_ARG;
    if (steps == 0) {
        if (rep) std::memset(rep, 0, sizeof(*rep));
        return PERKS_OK;
    }
    J2dSpec sp{a, b, wy, wx, steps, 0};
    return j2d_run(dtype, &sp, 1, o, (hipStream_t)stream, rep);
}

template <typename T>
int host2(int dtype, const T* h_in, int wy, int wx, T* h_out, int steps, const perks_stencil_opts* o,
          perks_run_report* rep)
{
    int rc = validate2(o, wy, wx, steps);
    if (rc != PERKS_OK) return rc;
    if (!h_in || !h_out) return PERKS_ERR_INVALID_ARG;
    const size_t bytes = (size_t)wy * wx * sizeof(T);
    T *d_a = nullptr, *d_b = nullptr;
    PERKS_HIP_TRY(hipMalloc(&d_a, bytes));
    if (hipMalloc(&d_b, bytes) != hipSuccess) {
        (void)hipFree(d_a);
        return PERKS_ERR_HIP;
    }
    perks_run_report local;
    perks_run_report* r = rep ? rep : &local;
    std::memset(r, 0, sizeof(*r));
    do {
        if (o->warmup) {
            const int ws = o->warmup_steps > 0 ? o->warmup_steps : 10;
            if (hipMemset(d_a, 0, bytes) != hipSuccess) { rc = PERKS_ERR_HIP; break; }
            rc = device2(dtype, d_a, d_b, wy, wx, ws, o, nullptr, r);
            if (rc != PERKS_OK) break;
        }
        if (hipMemcpy(d_a, h_in, bytes, hipMemcpyHostToDevice) != hipSuccess) { rc = PERKS_ERR_HIP; break; }
        rc = device2(dtype, d_a, d_b, wy, wx, steps, o, nullptr, r);
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

int perks_jacobi2d_iterative_f64(const double* h_in, int width_y, int width_x, double* h_out, int steps,
                                 const perks_stencil_opts* opts, perks_run_report* report)
{
    return perks::host2<double>(PERKS_F64, h_in, width_y, width_x, h_out, steps, opts, report);
}

int perks_jacobi2d_iterative_f32(const float* h_in, int width_y, int width_x, float* h_out, int steps,
                                 const perks_stencil_opts* opts, perks_run_report* report)
{
    return perks::host2<float>(PERKS_F32, h_in, width_y, width_x, h_out, steps, opts, report);
}

int perks_jacobi2d_device(int dtype, void* d_a, void* d_b, int width_y, int width_x, int steps,
                          const perks_stencil_opts* opts, void* stream, perks_run_report* report)
{
    return perks::device2(dtype, d_a, d_b, width_y, width_x, steps, opts, stream, report);
}

int perks_jacobi2d_batch(int dtype, perks_grid2d* grids, int ngrids, const perks_stencil_opts* opts, void* stream,
                         perks_run_report* report)
{
    if (!grids || ngrids <= 0 || !opts) return PERKS_ERR_INVALID_ARG;
    if (dtype < PERKS_F64 || dtype > PERKS_BF16) return PERKS_ERR_INVALID_ARG;
    std::vector<perks::J2dSpec> specs((size_t)ngrids);
    for (int i = 0; i < ngrids; ++i) {
        const int rc = perks::validate2(opts, grids[i].width_y, grids[i].width_x, grids[i].steps);
        if (rc != PERKS_OK) return rc;
        specs[i] = perks::J2dSpec{grids[i].a, grids[i].b, grids[i].width_y, grids[i].width_x, grids[i].steps, 0};
    }
    const int rc = perks::j2d_run(dtype, specs.data(), ngrids, opts, (hipStream_t)stream, report);
    for (int i = 0; i < ngrids; ++i) grids[i].result_in_b = specs[i].result_in_b;
    return rc;
}

int perks_jacobi2d_batch_noalloc(int dtype, perks_grid2d* grids, int ngrids, const perks_stencil_opts* opts,
                                 void* stream, perks_run_report* report)
{
    if (!grids || ngrids <= 0 || !opts) return PERKS_ERR_INVALID_ARG;
    if (dtype < PERKS_F64 || dtype > PERKS_BF16) return PERKS_ERR_INVALID_ARG;
    if (ngrids > PERKS_NOALLOC_MAX_GRIDS) {
        std::fprintf(stderr, "perks_amd: perks_jacobi2d_batch_noalloc: %d grids exceed its fixed scratch (at most %d)\n",
                     ngrids, PERKS_NOALLOC_MAX_GRIDS);
        return PERKS_ERR_UNSUPPORTED_SIZE;
    }
    std::vector<perks::J2dSpec> specs((size_t)ngrids);
    for (int i = 0; i < ngrids; ++i) {
        const int rc = perks::validate2(opts, grids[i].width_y, grids[i].width_x, grids[i].steps);
        if (rc != PERKS_OK) return rc;
        specs[i] = perks::J2dSpec{grids[i].a, grids[i].b, grids[i].width_y, grids[i].width_x, grids[i].steps, 0};
    }
    const int rc = perks::j2d_run_noalloc(dtype, specs.data(), ngrids, opts, (hipStream_t)stream, report);
    for (int i = 0; i < ngrids; ++i) grids[i].result_in_b = specs[i].result_in_b;
    return rc;
}

}  // extern "C"
