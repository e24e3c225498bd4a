// cg_driver.cpp -- the reference's CG driver (conjugateGradient/cg_driver.cu:53-428,
// myTest<ValueT,OffsetT,baseline,cacheMatrix,cacheVector>) on this library:
// matrix input, rhs = 1 and x = 0, one persistent CG solve, the reference's
// report and its optional host check.  The C++ template with the reference's
// signature is in include/perks_amd/cg_driver.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "api_common.h"
#include "cg.h"

namespace perks {
namespace {

struct DevBufs {
    std::vector<void*> ptrs;
    ~DevBufs()
    {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    int alloc(T** p, size_t n)
    {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return PERKS_ERR_HIP;
        ptrs.push_back(q);
        *p = (T*)q;
        return PERKS_OK;
    }
};

template <typename V>
int run_driver(const perks_cg_driver_opts* o, perks_cg_driver_result* res)
{
    const bool file = o->mtx_path && o->mtx_path[0];
    int32_t n = 0;
    int64_t nnz = 0;
    std::vector<int32_t> rowptr, col;
    std::vector<V> val;
    if (file) {
        int32_t *rp = nullptr, *cl = nullptr;
        double* vl = nullptr;
        const int rc = perks_mm_read_csr(o->mtx_path, &n, &nnz, &rp, &cl, &vl);
        if (rc != PERKS_OK) return rc;
        rowptr.assign(rp, rp + n + 1);
        col.assign(cl, cl + nnz);
        val.resize((size_t)nnz);
        for (int64_t i = 0; i < nnz; ++i) val[i] = (V)vl[i];
        perks_free_host(rp);
        perks_free_host(cl);
        perks_free_host(vl);
    } else {
        n = o->vdata_n > 0 ? o->vdata_n : 1048576;
        if (n < 2) return PERKS_ERR_INVALID_ARG;
        nnz = (int64_t)(n - 2) * 3 + 4;
        rowptr.resize((size_t)n + 1);
        col.resize((size_t)nnz);
        val.resize((size_t)nnz);
        const int rc = sizeof(V) == 8 ? perks_gen_tridiag_f64(n, rowptr.data(), col.data(), (double*)val.data())
                                      : perks_gen_tridiag_f32(n, rowptr.data(), col.data(), (float*)val.data());
        if (rc != PERKS_OK) return rc;
    }
    const int max_iter = o->max_iter > 0 ? o->max_iter : (file ? 10 * n : n);

    DevBufs d;
    int32_t *d_rp = nullptr, *d_col = nullptr;
    V *d_val = nullptr, *d_x = nullptr, *d_b = nullptr;
    int rc = d.alloc(&d_rp, (size_t)n + 1);
    if (rc == PERKS_OK) rc = d.alloc(&d_col, (size_t)nnz);
    if (rc == PERKS_OK) rc = d.alloc(&d_val, (size_t)nnz);
    if (rc == PERKS_OK) rc = d.alloc(&d_x, (size_t)n);
    if (rc == PERKS_OK) rc = d.alloc(&d_b, (size_t)n);
    if (rc != PERKS_OK) return rc;
    const std::vector<V> ones((size_t)n, (V)1), zeros((size_t)n, (V)0);
    PERKS_HIP_TRY(hipMemcpy(d_rp, rowptr.data(), sizeof(int32_t) * rowptr.size(), hipMemcpyHostToDevice));
    PERKS_HIP_TRY(hipMemcpy(d_col, col.data(), sizeof(int32_t) * col.size(), hipMemcpyHostToDevice));
    PERKS_HIP_TRY(hipMemcpy(d_val, val.data(), sizeof(V) * val.size(), hipMemcpyHostToDevice));
    PERKS_HIP_TRY(hipMemcpy(d_b, ones.data(), sizeof(V) * n, hipMemcpyHostToDevice));

    perks_cg_opts co;
    perks_cg_opts_init(&co);
    co.tol = o->tol > 0 ? o->tol : 1e-10;
    co.relative = 0;  // the reference stops at r.r < tol^2 (cg_kernels.cuh:834)
    co.static_iter = o->static_iter;
    co.x0_is_zero = 1;
    std::vector<double> hist((size_t)max_iter + 2);
    perks_cg_report rep;
    auto solve = [&](int iters) -> int {
        PERKS_HIP_TRY(hipMemcpy(d_x, zeros.data(), sizeof(V) * n, hipMemcpyHostToDevice));
        co.max_iter = iters;
        std::memset(&rep, 0, sizeof(rep));
        rep.history = hist.data();
        rep.hist_cap = (int32_t)hist.size();
        if (sizeof(V) == 8) {
            perks_csr A;
            std::memset(&A, 0, sizeof(A));
            A.n = n;
            A.nnz = nnz;
            A.rowptr = d_rp;
            A.col = d_col;
            A.val = (const double*)d_val;
            return perks_cg_solve_csr(&A, (const double*)d_b, (double*)d_x, &co, nullptr, &rep);
        }
        perks_csr_f32 A;
        std::memset(&A, 0, sizeof(A));
        A.n = n;
        A.nnz = nnz;
        A.rowptr = d_rp;
        A.col = d_col;
        A.val = (const float*)d_val;
        return perks_cg_solve_csr_f32(&A, (const float*)d_b, (float*)d_x, &co, nullptr, &rep);
    };
    if (o->warmup) {
        rc = solve(2);
        if (rc != PERKS_OK && rc != PERKS_ERR_NOT_CONVERGED) return rc;
    }
    rc = solve(max_iter);
    if (rc != PERKS_OK && rc != PERKS_ERR_NOT_CONVERGED) return rc;

    std::memset(res, 0, sizeof(*res));
    res->n = n;
    res->nnz = nnz;
    res->iterations = rep.iterations;
    res->residual = std::sqrt(rep.final_rr);
    res->elapsed_ms = rep.elapsed_ms;
    res->ms_per_iter = rep.iterations > 0 ? rep.elapsed_ms / rep.iterations : 0.0;
    // The reference's byte model per iteration (cg_driver.cu:359-361), over the iterations run.
    const double spmv = (double)nnz * sizeof(V) + (double)(nnz + n + 1) * sizeof(int32_t);
    const double bytes = (double)rep.iterations * (spmv + 6.0 * n * sizeof(V));
    res->gbps = rep.elapsed_ms > 0 ? bytes / (rep.elapsed_ms * 1e-3) / 1e9 : 0.0;
    res->grid = rep.grid;
    res->code = rc;
    res->check_error = -1.0;
    if (o->check) {
        std::vector<V> x((size_t)n);
        PERKS_HIP_TRY(hipMemcpy(x.data(), d_x, sizeof(V) * n, hipMemcpyDeviceToHost));
        V err = 0;
        for (int32_t i = 0; i < n; ++i) {
            V rsum = 0;
            for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) rsum += val[k] * x[col[k]];
            const V diff = std::fabs(rsum - (V)1);
            if (diff > err) err = diff;
        }
        res->check_error = (double)err;
    }
    if (o->verbose) {
        std::printf("GPU Final, residual = %e, kernel execution time = %f ms in %d iteration\n", res->residual,
                    res->elapsed_ms, res->iterations);
        std::printf("max iter is %d\n", max_iter);
        std::printf("%s\t", file ? o->mtx_path : "virtualdata");
        std::printf("<%f ms: %fGB/s>\n", res->ms_per_iter, res->gbps);
        if (o->check) std::printf("* Test Summary:  Error amount = %f \n", res->check_error);
    }
    return PERKS_OK;
}

}  // namespace
}  // namespace perks

extern "C" int perks_cg_driver(const perks_cg_driver_opts* opts, perks_cg_driver_result* result)
{
    if (!opts || !result) return PERKS_ERR_INVALID_ARG;
    if (opts->value_bytes == 8) return perks::run_driver<double>(opts, result);
    if (opts->value_bytes == 4) return perks::run_driver<float>(opts, result);
    return PERKS_ERR_INVALID_ARG;
}
