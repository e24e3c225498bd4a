// cg.hip -- persistent conjugate gradient (CSR and DIA; FP64, and FP32 inner
// iterations inside an FP64 iterative-refinement loop).
//
// Replaces the reference's gpuConjugateGradient_cub (conjugateGradient/
// cg_kernels.cuh:620-986) and its dispatcher (cg_dispatcher.cuh:559-916): one
// cooperative launch runs the whole solve.  The order of operations is the
// reference's (cg_kernels.cuh:721-975):
//     r = b + (-1)*A x ;  r1 = r.r
//     while !(r1 < thr):  p = 1*r + beta*p (beta = r1/r0 from the 2nd iteration)
//                         Ap = A p ; alpha = r1 / p.Ap
//                         x = alpha*p + x ; r = (-alpha)*Ap + r ; r0 = r1 ; r1 = r.r
// with two changes made for MI355X and for reproducibility:
//   * two grid barriers per iteration instead of five: the p update is fused
//     into the SpMV phase (a row's p is written by its owner, and neighbours
//     re-form the same value r_j + beta*p_j from the previous p buffer, which
//     is ping-ponged), and the x/r update is fused with the r.r partials;
//   * dot products are reduced in a fixed order instead of with atomics: the
//     vector is cut into 4096-element segments, lane l of a 256-thread
//     workgroup sums elements l, l+256, ..., l+15*256 of a segment in order,
//     the 256 lane sums are folded pairwise with strides 128..1, and every
//     workgroup folds all segment sums the same way (no extra barrier; all
//     workgroups obtain the identical value).  The CPU oracle restates this
//     order, so the residual history is reproducible bit for bit.
// Built with -ffp-contract=off: every multiply and add rounds on its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "api_common.h"
#include "cg.h"
#include "grid_sync.h"

namespace perks {

namespace {

constexpr int kLanes = 256;
constexpr int kPerLane = 16;
constexpr int kSeg = kLanes * kPerLane;
constexpr int kMaxDiag = 16;
constexpr int kCsrChunk = 8;  // row entries loaded per predicated chunk
constexpr int kMaxWgPerCu = kCoopMaxWavesPerCu / 4;  // 256-thread workgroups

// Matrix formats the kernel is built for.  DIA7 is the 7-diagonal form with
// the diagonal count fixed at compile time (every load of a row issued at once).
enum { FMT_CSR = 0, FMT_DIA = 1, FMT_DIA7 = 2 };

template <typename V>
struct CgArgs {
    int n;
    int nseg;
    // matrix (values in V)
    const int32_t* rowptr;
    const int32_t* col;
    const V* val;
    int ndiag;
    int32_t offs[kMaxDiag];
    // vectors
    V* x;
    V* r;
    V* p0;
    V* p1;
    V* ap;
    const V* b;
    // reduction scratch: two partial arrays of nseg doubles each
    double* part_a;
    double* part_b;
    // control
    double tol2;        // tol^2
    int relative;       // threshold relative to b.b
    int max_iter;
    double* hist;       // r.r after each iteration (hist[0] = initial), may be null
    int hist_cap;
    int* iters_out;
    double* rr_out;
    double* bb_out;
    GridSync* sync;
    unsigned long long timeout;
    int x_is_zero;      // skip the initial SpMV
    int static_iter;    // run exactly max_iter iterations
};

// Folds the 256 per-lane values of a workgroup pairwise with strides
// 128, 64, ..., 1 (v[l] += v[l + s]); returns the result in every lane.
__device__ __forceinline__ double fold256(double v, double* lds)
{
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    if (t < 128) lds[t] = lds[t] + lds[t + 128];
    __syncthreads();
    if (t < 64) {
        double u = lds[t] + lds[t + 64];
        // strides 32..1 inside wave 0
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) u = u + __shfl_down(u, s, 64);
        if (t == 0) lds[0] = u;
    }
    __syncthreads();
    const double r = lds[0];
    __syncthreads();
    return r;
}

// Sum of the segment partials in the fixed order: lane l sums partials
// l, l+256, ... in order, then the 256-lane fold.
__device__ __forceinline__ double reduce_parts(const double* part, int nseg, double* lds)
{
    double acc = 0.0;
    for (int j = threadIdx.x; j < nseg; j += kLanes) acc = acc + part[j];
    return fold256(acc, lds);
}

// Vector entry j of the operand: r_j for the first iteration (p_1 = r_0),
// else p_j = r_j + beta * pold_j formed on the fly.
template <typename V, bool FIRST>
__device__ __forceinline__ V operand(const V* __restrict__ r, const V* __restrict__ pold, V beta, int j)
{
    return FIRST ? r[j] : r[j] + beta * pold[j];
}

// Row i of A times the operand.  Entries are summed in storage order from
// zero; DIA entries whose column falls outside [0, n) are skipped.  All loads
// are unconditional (out-of-range columns read row i itself and the sum keeps
// its value through a select), so a row's loads are in flight together.
template <typename V, int FMT, bool FIRST>
__device__ __forceinline__ V row_product(const CgArgs<V>& a, int i, const V* __restrict__ r,
                                         const V* __restrict__ pold, V beta)
{
    V s = (V)0;
    if (FMT == FMT_CSR) {
        const int k0 = a.rowptr[i], k1 = a.rowptr[i + 1];
        for (int kb = k0; kb < k1; kb += kCsrChunk) {
            int cj[kCsrChunk];
            V cv[kCsrChunk];
#pragma unroll
            for (int u = 0; u < kCsrChunk; ++u) {
                const int k = (kb + u < k1) ? kb + u : k0;
                cj[u] = a.col[k];
                cv[u] = a.val[k];
            }
#pragma unroll
            for (int u = 0; u < kCsrChunk; ++u) {
                const V prod = cv[u] * operand<V, FIRST>(r, pold, beta, cj[u]);
                s = (kb + u < k1) ? s + prod : s;
            }
        }
    } else {
        const int nd = FMT == FMT_DIA7 ? 7 : a.ndiag;
#pragma unroll
        for (int d = 0; d < (FMT == FMT_DIA7 ? 7 : kMaxDiag); ++d) {
            if (FMT != FMT_DIA7 && d >= nd) break;
            const long long j = (long long)i + a.offs[d];
            const bool in = (unsigned long long)j < (unsigned long long)a.n;
            const int jc = in ? (int)j : i;
            const V prod = a.val[(size_t)d * a.n + i] * operand<V, FIRST>(r, pold, beta, jc);
            s = in ? s + prod : s;
        }
    }
    return s;
}

template <typename V>
__device__ __forceinline__ double prod_d(V a, V b)
{
    return (double)a * (double)b;  // exact for float operands; the product itself for double
}

// Phase 1 of an iteration over this workgroup's segments: p = r + beta*p
// (stored into pnew), Ap = A p, partial sums of p.Ap into part_a.
template <typename V, int FMT, bool FIRST>
__device__ __forceinline__ void phase_spmv(const CgArgs<V>& a, const V* pold, V* pnew, V beta, double* lds)
{
    const int t = threadIdx.x;
    for (int s = blockIdx.x; s < a.nseg; s += gridDim.x) {
        double acc = 0.0;
#pragma unroll 2
        for (int kk = 0; kk < kPerLane; ++kk) {
            const int i = s * kSeg + kk * kLanes + t;
            if (i < a.n) {
                const V pi = operand<V, FIRST>(a.r, pold, beta, i);
                const V api = row_product<V, FMT, FIRST>(a, i, a.r, pold, beta);
                pnew[i] = pi;
                a.ap[i] = api;
                acc = acc + prod_d(pi, api);
            }
        }
        const double ps = fold256(acc, lds);
        if (t == 0) a.part_a[s] = ps;
    }
}

template <typename V, int FMT>
__global__ void __launch_bounds__(kLanes) cg_persistent_kernel(CgArgs<V> a)
{
    __shared__ double lds[kLanes];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    const int me = blockIdx.x;

    // ---- r = b + (-1) * A x ; partials of r.r and b.b ----
    for (int s = me; s < a.nseg; s += G) {
        double rr = 0.0, bb = 0.0;
#pragma unroll 2
        for (int k = 0; k < kPerLane; ++k) {
            const int i = s * kSeg + k * kLanes + t;
            if (i < a.n) {
                const V bi = a.b[i];
                const V ri = a.x_is_zero ? bi : bi + (V)(-1) * row_product<V, FMT, true>(a, i, a.x, a.x, (V)0);
                a.r[i] = ri;
                rr = rr + prod_d(ri, ri);
                bb = bb + prod_d(bi, bi);
            }
        }
        const double prr = fold256(rr, lds);
        const double pbb = fold256(bb, lds);
        if (t == 0) {
            a.part_a[s] = prr;
            a.part_b[s] = pbb;
        }
    }
    if (!grid_sync(a.sync, G, a.timeout)) return;
    double r1 = reduce_parts(a.part_a, a.nseg, lds);
    const double bbv = reduce_parts(a.part_b, a.nseg, lds);
    const double thr = a.relative ? a.tol2 * bbv : a.tol2;
    if (me == 0 && t == 0) {
        if (a.hist && a.hist_cap > 0) a.hist[0] = r1;
        *a.bb_out = bbv;
    }
    if (!grid_sync(a.sync, G, a.timeout)) return;  // part_a/part_b are rewritten below

    double r0 = 0.0;
    int k = 1;
    V* pold = a.p0;
    V* pnew = a.p1;
    while ((a.static_iter || !(r1 < thr)) && k <= a.max_iter) {
        // ---- p = r + beta*p ; Ap = A p ; partials of p.Ap ----
        if (k == 1) {
            phase_spmv<V, FMT, true>(a, pold, pnew, (V)0, lds);
        } else {
            phase_spmv<V, FMT, false>(a, pold, pnew, (V)(r1 / r0), lds);
        }
        if (!grid_sync(a.sync, G, a.timeout)) return;
        const double pap = reduce_parts(a.part_a, a.nseg, lds);
        const double alpha_d = r1 / pap;
        const V alpha = (V)alpha_d;
        const V na = -alpha;
        // ---- x = alpha*p + x ; r = (-alpha)*Ap + r ; partials of r.r ----
        for (int s = me; s < a.nseg; s += G) {
            double acc = 0.0;
#pragma unroll 4
            for (int kk = 0; kk < kPerLane; ++kk) {
                const int i = s * kSeg + kk * kLanes + t;
                if (i < a.n) {
                    a.x[i] = alpha * pnew[i] + a.x[i];
                    const V ri = na * a.ap[i] + a.r[i];
                    a.r[i] = ri;
                    acc = acc + prod_d(ri, ri);
                }
            }
            const double ps = fold256(acc, lds);
            if (t == 0) a.part_b[s] = ps;
        }
        if (!grid_sync(a.sync, G, a.timeout)) return;
        r0 = r1;
        r1 = reduce_parts(a.part_b, a.nseg, lds);
        if (me == 0 && t == 0 && a.hist && k < a.hist_cap) a.hist[k] = r1;
        ++k;
        V* tmp = pold;
        pold = pnew;
        pnew = tmp;
    }
    if (me == 0 && t == 0) {
        *a.iters_out = k - 1;
        *a.rr_out = r1;
    }
}

// ---- elementwise helpers for iterative refinement ------------------------
__global__ void to_f32(const double* __restrict__ s, float* __restrict__ d, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = (float)s[i];
}

__global__ void add_f32_to_f64(double* __restrict__ x, const float* __restrict__ d, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = x[i] + (double)d[i];
}

__global__ void zero_f32(float* __restrict__ d, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = 0.0f;
}

// ---- scratch ---------------------------------------------------------------
struct CgScratch {
    size_t n = 0;
    int nseg = 0;
    void* p0 = nullptr;  // sized for double
    void* p1 = nullptr;
    void* ap = nullptr;
    void* r = nullptr;
    float* rf = nullptr;   // FP32 right-hand side / residual (refinement)
    float* xf = nullptr;   // FP32 correction (refinement)
    float* valf = nullptr; // FP32 copy of the matrix values (refinement)
    size_t nvalf = 0;
    double* parts = nullptr;
    double* small = nullptr;  // [0] rr, [1] bb ; int at [2]
    double* hist = nullptr;
    int hist_cap = 0;
};
std::mutex g_cg_mu;
CgScratch g_cg[64];

// Replaces *slot with a fresh allocation of `bytes`.  The slot is nulled
// before the allocation, so a failed allocation leaves nothing freed behind.
int realloc_dev(void** slot, size_t bytes)
{
    if (*slot) {
        void* old = *slot;
        *slot = nullptr;
        PERKS_HIP_TRY(hipFree(old));
    }
    PERKS_HIP_TRY(hipMalloc(slot, bytes));
    return PERKS_OK;
}

#define PERKS_TRY(expr)                 \
    do {                                \
        const int perks_rc_ = (expr);   \
        if (perks_rc_ != PERKS_OK) return perks_rc_; \
    } while (0)

int ensure(CgScratch& s, size_t n, size_t nval32, int hist_cap)
{
    if (s.n < n) {
        // The sizes are recorded only once every buffer is in place.
        s.n = 0;
        s.nseg = 0;
        const int nseg = (int)((n + kSeg - 1) / kSeg);
        PERKS_TRY(realloc_dev(&s.p0, n * 8));
        PERKS_TRY(realloc_dev(&s.p1, n * 8));
        PERKS_TRY(realloc_dev(&s.ap, n * 8));
        PERKS_TRY(realloc_dev(&s.r, n * 8));
        PERKS_TRY(realloc_dev((void**)&s.rf, n * 4));
        PERKS_TRY(realloc_dev((void**)&s.xf, n * 4));
        PERKS_TRY(realloc_dev((void**)&s.parts, sizeof(double) * 2 * (size_t)nseg));
        s.n = n;
        s.nseg = nseg;
    }
    if (nval32 > s.nvalf) {
        s.nvalf = 0;
        PERKS_TRY(realloc_dev((void**)&s.valf, nval32 * 4));
        s.nvalf = nval32;
    }
    if (!s.small) PERKS_HIP_TRY(hipMalloc(&s.small, 64));
    if (s.hist_cap < hist_cap) {
        s.hist_cap = 0;
        PERKS_TRY(realloc_dev((void**)&s.hist, sizeof(double) * (size_t)hist_cap));
        s.hist_cap = hist_cap;
    }
    return PERKS_OK;
}

template <typename V, int FMT>
int launch_cg(CgArgs<V>& a, GridSync* gs, hipStream_t stream, int blkpsm, int* grid_out)
{
    const void* fn = reinterpret_cast<const void*>(&cg_persistent_kernel<V, FMT>);
    int occ = 0;
    PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kLanes, 0));
    if (occ <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
    if (blkpsm > 0 && blkpsm < occ) occ = blkpsm;
    // At most half the wave slots: 4 workgroups of 4 waves per CU keep enough
    // loads in flight, and the cooperative grid never needs every slot free.
    if (occ > kMaxWgPerCu) occ = kMaxWgPerCu;
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    int G = caps.cu_count * occ;
    if (G > a.nseg) G = a.nseg > 0 ? a.nseg : 1;
    a.sync = gs;
    a.timeout = grid_sync_timeout();
    void* args[] = {&a};
    PERKS_HIP_TRY(launch_persistent(fn, dim3(G), dim3(kLanes), args, 0, stream));
    if (grid_out) *grid_out = G;
    return PERKS_OK;
}

template <typename V>
int launch_cg_fmt(int fmt, CgArgs<V>& a, GridSync* gs, hipStream_t stream, int blkpsm, int* grid_out)
{
    if (fmt == FMT_CSR) return launch_cg<V, FMT_CSR>(a, gs, stream, blkpsm, grid_out);
    if (a.ndiag == 7) return launch_cg<V, FMT_DIA7>(a, gs, stream, blkpsm, grid_out);
    return launch_cg<V, FMT_DIA>(a, gs, stream, blkpsm, grid_out);
}

}  // namespace

// Solves A x = b.  Host-visible results go to `rep`.
int cg_solve(int fmt, const perks_csr* csr, const perks_dia* dia, const double* d_b, double* d_x,
             const perks_cg_opts* o, void* stream_v, perks_cg_report* rep)
{
    hipStream_t stream = (hipStream_t)stream_v;
    if (!o || !d_b || !d_x || !rep) return PERKS_ERR_INVALID_ARG;
    const int n = fmt == FMT_CSR ? (csr ? csr->n : -1) : (dia ? dia->n : -1);
    if (n <= 0) return PERKS_ERR_INVALID_ARG;
    if (fmt == FMT_DIA && (dia->ndiag <= 0 || dia->ndiag > 16)) return PERKS_ERR_UNSUPPORTED;
    if (o->max_iter < 0) return PERKS_ERR_INVALID_ARG;
    const size_t nval = fmt == FMT_CSR ? (size_t)csr->nnz : (size_t)dia->ndiag * (size_t)n;
    const bool refine = o->precision == PERKS_CG_F32_REFINE;
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return PERKS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_cg_mu);
    CgScratch& S = g_cg[dev];
    const int hist_cap = std::max(1, o->max_iter + 1);
    int rc = ensure(S, (size_t)n, refine ? nval : 0, hist_cap);
    if (rc != PERKS_OK) return rc;
    SyncLease lease;
    rc = lease.acquire();
    if (rc != PERKS_OK) return rc;

    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    int grid = 0;
    std::vector<double> hist_host;
    int total_iters = 0, outer = 0;
    double rr = 0.0, bb = 0.0;

    auto fill64 = [&](CgArgs<double>& a) {
        std::memset(&a, 0, sizeof(a));
        a.n = n;
        a.nseg = S.nseg;
        if (fmt == FMT_CSR) {
            a.rowptr = csr->rowptr;
            a.col = csr->col;
            a.val = csr->val;
        } else {
            a.val = dia->val;
            a.ndiag = dia->ndiag;
            for (int d = 0; d < dia->ndiag; ++d) a.offs[d] = dia->offsets[d];
        }
        a.part_a = S.parts;
        a.part_b = S.parts + S.nseg;
        a.rr_out = S.small;
        a.bb_out = S.small + 1;
        a.iters_out = (int*)(S.small + 2);
    };

    if (!refine) {
        CgArgs<double> a;
        fill64(a);
        a.x = d_x;
        a.r = (double*)S.r;
        a.p0 = (double*)S.p0;
        a.p1 = (double*)S.p1;
        a.ap = (double*)S.ap;
        a.b = d_b;
        a.tol2 = o->tol * o->tol;
        a.relative = o->relative;
        a.max_iter = o->max_iter;
        a.hist = S.hist;
        a.hist_cap = hist_cap;
        a.x_is_zero = o->x0_is_zero;
        a.static_iter = o->static_iter;
        rc = launch_cg_fmt<double>(fmt, a, lease.gs, stream, o->blkpsm, &grid);
        if (rc != PERKS_OK) return rc;
    } else {
        // FP32 copy of the matrix values, once per solve.
        const int nb = (int)std::min<size_t>((nval + 255) / 256, 65536);
        const double* v64 = fmt == FMT_CSR ? csr->val : dia->val;
        // (values are converted with the same kernel as vectors)
        hipLaunchKernelGGL(to_f32, dim3(nb), dim3(256), 0, stream, v64, S.valf, (int)nval);
        PERKS_HIP_TRY(hipGetLastError());
    }
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    GridSync* gs = lease.gs;
    if (!refine) {
        rc = grid_sync_check(gs, stream);
        if (rc != PERKS_OK) return rc;
    }

    if (!refine) {
        int it = 0;
        double small[3];
        PERKS_HIP_TRY(hipMemcpy(small, S.small, sizeof(small), hipMemcpyDeviceToHost));
        std::memcpy(&it, &small[2], sizeof(int));
        rr = small[0];
        bb = small[1];
        total_iters = it;
        hist_host.resize((size_t)it + 1);
        PERKS_HIP_TRY(hipMemcpy(hist_host.data(), S.hist, sizeof(double) * hist_host.size(), hipMemcpyDeviceToHost));
    } else {
        // Iterative refinement: r = b - A x in FP64, an FP32 CG on A d = r to a
        // relative tolerance (inner_tol, relaxed on the last steps to what the
        // outer target still needs), x += d, until the FP64 residual meets tol.
        const int nb = (int)std::min<size_t>(((size_t)n + 255) / 256, 65536);
        float elapsed = ms;
        double rr_prev = -1.0;
        const int max_outer = o->max_outer > 0 ? o->max_outer : 20;
        for (outer = 0; outer <= max_outer; ++outer) {
            // (1) FP64 residual with zero iterations: the kernel's initial phase.
            CgArgs<double> a;
            fill64(a);
            a.x = d_x;
            a.r = (double*)S.r;
            a.p0 = (double*)S.p0;
            a.p1 = (double*)S.p1;
            a.ap = (double*)S.ap;
            a.b = d_b;
            a.tol2 = o->tol * o->tol;
            a.relative = o->relative;
            a.max_iter = 0;
            a.hist = nullptr;
            a.x_is_zero = (outer == 0) ? o->x0_is_zero : 0;
            PERKS_HIP_TRY(hipEventRecord(ev0, stream));
            rc = launch_cg_fmt<double>(fmt, a, lease.gs, stream, o->blkpsm, &grid);
            if (rc != PERKS_OK) return rc;
            double small[3];
            PERKS_HIP_TRY(hipMemcpyAsync(small, S.small, sizeof(small), hipMemcpyDeviceToHost, stream));
            PERKS_HIP_TRY(hipStreamSynchronize(stream));
            rc = grid_sync_check(gs, stream);
            if (rc != PERKS_OK) return rc;
            rr = small[0];
            bb = small[1];
            hist_host.push_back(rr);
            const double thr = o->relative ? o->tol * o->tol * bb : o->tol * o->tol;
            if (rr < thr || outer == max_outer || rr == rr_prev) {
                PERKS_HIP_TRY(hipEventRecord(ev1, stream));
                PERKS_HIP_TRY(hipEventSynchronize(ev1));
                float m2 = 0.f;
                PERKS_HIP_TRY(hipEventElapsedTime(&m2, ev0, ev1));
                elapsed += m2;
                break;
            }
            rr_prev = rr;
            // (2) FP32 inner CG on A d = r, d0 = 0.
            hipLaunchKernelGGL(to_f32, dim3(nb), dim3(256), 0, stream, (const double*)S.r, S.rf, n);
            hipLaunchKernelGGL(zero_f32, dim3(nb), dim3(256), 0, stream, S.xf, n);
            PERKS_HIP_TRY(hipGetLastError());
            CgArgs<float> f;
            std::memset(&f, 0, sizeof(f));
            f.n = n;
            f.nseg = S.nseg;
            if (fmt == FMT_CSR) {
                f.rowptr = csr->rowptr;
                f.col = csr->col;
            } else {
                f.ndiag = dia->ndiag;
                for (int d = 0; d < dia->ndiag; ++d) f.offs[d] = dia->offsets[d];
            }
            f.val = S.valf;
            f.x = S.xf;
            f.r = (float*)S.r;  // the FP64 residual buffer is free now: reuse it for FP32 vectors
            f.p0 = (float*)S.p0;
            f.p1 = (float*)S.p1;
            f.ap = (float*)S.ap;
            f.b = S.rf;
            // The inner solve stops at relative tolerance max(inner_tol, sqrt(thr / rr) / 2):
            // never further than the outer target needs.
            f.tol2 = o->inner_tol * o->inner_tol;
            if (rr > 0.0 && 0.25 * thr / rr > f.tol2) f.tol2 = 0.25 * thr / rr;
            f.relative = 1;
            f.max_iter = o->max_inner > 0 ? o->max_inner : o->max_iter;
            f.part_a = S.parts;
            f.part_b = S.parts + S.nseg;
            f.rr_out = S.small;
            f.bb_out = S.small + 1;
            f.iters_out = (int*)(S.small + 2);
            f.x_is_zero = 1;
            rc = launch_cg_fmt<float>(fmt, f, lease.gs, stream, o->blkpsm, &grid);
            if (rc != PERKS_OK) return rc;
            // (3) x += d
            hipLaunchKernelGGL(add_f32_to_f64, dim3(nb), dim3(256), 0, stream, d_x, S.xf, n);
            PERKS_HIP_TRY(hipGetLastError());
            PERKS_HIP_TRY(hipEventRecord(ev1, stream));
            PERKS_HIP_TRY(hipEventSynchronize(ev1));
            float m2 = 0.f;
            PERKS_HIP_TRY(hipEventElapsedTime(&m2, ev0, ev1));
            elapsed += m2;
            rc = grid_sync_check(gs, stream);
            if (rc != PERKS_OK) return rc;
            PERKS_HIP_TRY(hipMemcpy(small, S.small, sizeof(small), hipMemcpyDeviceToHost));
            int it = 0;
            std::memcpy(&it, &small[2], sizeof(int));
            total_iters += it;
        }
        ms = elapsed;
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);

    rep->iterations = total_iters;
    rep->outer_iterations = refine ? outer : 0;
    rep->final_rr = rr;
    rep->bb = bb;
    rep->elapsed_ms = ms;
    rep->grid = grid;
    const int ncopy = std::min<int>((int)hist_host.size(), rep->hist_cap);
    if (rep->history && ncopy > 0) std::memcpy(rep->history, hist_host.data(), sizeof(double) * (size_t)ncopy);
    rep->hist_len = (int)hist_host.size();
    const double thr = o->relative ? o->tol * o->tol * bb : o->tol * o->tol;
    return (rr < thr || (!refine && o->static_iter)) ? PERKS_OK : PERKS_ERR_NOT_CONVERGED;
}

// FP32 CG on FP32 data (the reference's myTest<float,...>).
int cg_solve_f32(int fmt, const perks_csr_f32* csr, const perks_dia_f32* dia, const float* d_b, float* d_x,
                 const perks_cg_opts* o, void* stream_v, perks_cg_report* rep)
{
    hipStream_t stream = (hipStream_t)stream_v;
    if (!o || !d_b || !d_x || !rep) return PERKS_ERR_INVALID_ARG;
    const int n = fmt == FMT_CSR ? (csr ? csr->n : -1) : (dia ? dia->n : -1);
    if (n <= 0 || o->max_iter < 0) return PERKS_ERR_INVALID_ARG;
    if (fmt == FMT_DIA && (dia->ndiag <= 0 || dia->ndiag > kMaxDiag)) return PERKS_ERR_UNSUPPORTED;
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return PERKS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_cg_mu);
    CgScratch& S = g_cg[dev];
    const int hist_cap = std::max(1, o->max_iter + 1);
    int rc = ensure(S, (size_t)n, 0, hist_cap);
    if (rc != PERKS_OK) return rc;
    SyncLease lease;
    rc = lease.acquire();
    if (rc != PERKS_OK) return rc;
    CgArgs<float> f;
    std::memset(&f, 0, sizeof(f));
    f.n = n;
    f.nseg = S.nseg;
    if (fmt == FMT_CSR) {
        f.rowptr = csr->rowptr;
        f.col = csr->col;
        f.val = csr->val;
    } else {
        f.ndiag = dia->ndiag;
        for (int d = 0; d < dia->ndiag; ++d) f.offs[d] = dia->offsets[d];
        f.val = dia->val;
    }
    f.x = d_x;
    f.r = (float*)S.r;
    f.p0 = (float*)S.p0;
    f.p1 = (float*)S.p1;
    f.ap = (float*)S.ap;
    f.b = d_b;
    const float tf = (float)o->tol;
    f.tol2 = (double)(tf * tf);
    f.relative = o->relative;
    f.max_iter = o->max_iter;
    f.hist = S.hist;
    f.hist_cap = hist_cap;
    f.part_a = S.parts;
    f.part_b = S.parts + S.nseg;
    f.rr_out = S.small;
    f.bb_out = S.small + 1;
    f.iters_out = (int*)(S.small + 2);
    f.x_is_zero = o->x0_is_zero;
    f.static_iter = o->static_iter;
    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    int grid = 0;
    rc = launch_cg_fmt<float>(fmt, f, lease.gs, stream, o->blkpsm, &grid);
    if (rc != PERKS_OK) return rc;
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    rc = grid_sync_check(lease.gs, stream);
    if (rc != PERKS_OK) return rc;
    double small[3];
    PERKS_HIP_TRY(hipMemcpy(small, S.small, sizeof(small), hipMemcpyDeviceToHost));
    int it = 0;
    std::memcpy(&it, &small[2], sizeof(int));
    std::vector<double> hist((size_t)it + 1);
    PERKS_HIP_TRY(hipMemcpy(hist.data(), S.hist, sizeof(double) * hist.size(), hipMemcpyDeviceToHost));
    rep->iterations = it;
    rep->outer_iterations = 0;
    rep->final_rr = small[0];
    rep->bb = small[1];
    rep->elapsed_ms = ms;
    rep->grid = grid;
    const int ncopy = std::min<int>((int)hist.size(), rep->hist_cap);
    if (rep->history && ncopy > 0) std::memcpy(rep->history, hist.data(), sizeof(double) * (size_t)ncopy);
    rep->hist_len = (int)hist.size();
    const double thr = o->relative ? f.tol2 * small[1] : f.tol2;
    return (small[0] < thr || o->static_iter) ? PERKS_OK : PERKS_ERR_NOT_CONVERGED;
}

int release_cg_scratch()
{
    int dev = 0;
    PERKS_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_cg_mu);
    if (dev < 0 || dev >= 64) return PERKS_OK;
    CgScratch& s = g_cg[dev];
    for (void* p : {s.p0, s.p1, s.ap, s.r, (void*)s.rf, (void*)s.xf, (void*)s.valf, (void*)s.parts,
                    (void*)s.small, (void*)s.hist})
        if (p) PERKS_HIP_TRY(hipFree(p));
    s = CgScratch{};
    return PERKS_OK;
}

// ---- Poisson 7-point generators (on the device) ------------------------------
// The operator of conjugateGradient's 3D Poisson test problem; the CSR row
// order and ascending columns match what the reference's COO->CSR conversion
// produces (conjugateGradient/matrix/sparse_matrix.h:666-734).
namespace {

__global__ void poisson_csr_kernel(int nx, int ny, int nz, int32_t* rowptr, int32_t* col, double* val)
{
    const long long n = (long long)nx * ny * nz;
    const long long sxy = (long long)nx * ny;
    for (long long row = blockIdx.x * (long long)blockDim.x + threadIdx.x; row < n;
         row += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(row % nx), y = (int)((row / nx) % ny), z = (int)(row / sxy);
        // entries before this row: 7 per row minus the faces cut by the domain
        // for every earlier row (a closed form per (z, y, x)).
        long long k = 7 * row;
        // rows missing their -z neighbour: z == 0 -> first sxy rows
        k -= std::min<long long>(row, sxy);
        // rows missing their +z neighbour: z == nz-1 -> last sxy rows
        k -= std::max<long long>(0, row - (n - sxy));
        // -y / +y: per plane, first / last nx rows of each plane
        const long long plane_rows_before = (long long)z * sxy;
        k -= (long long)z * nx + std::min<long long>(row - plane_rows_before, nx);
        k -= (long long)z * nx + std::max<long long>(0, (row - plane_rows_before) - (sxy - nx));
        // -x / +x: one per line
        const long long lines_before = (long long)z * ny + y;
        k -= lines_before + (x > 0 ? 1 : 0);
        k -= lines_before;
        rowptr[row] = (int32_t)k;
        int e = 0;
        auto put = [&](long long c, double v) {
            col[k + e] = (int32_t)c;
            val[k + e] = v;
            ++e;
        };
        if (z > 0) put(row - sxy, -1.0);
        if (y > 0) put(row - nx, -1.0);
        if (x > 0) put(row - 1, -1.0);
        put(row, 6.0);
        if (x < nx - 1) put(row + 1, -1.0);
        if (y < ny - 1) put(row + nx, -1.0);
        if (z < nz - 1) put(row + sxy, -1.0);
        if (row == n - 1) rowptr[n] = (int32_t)(k + e);
    }
}

__global__ void poisson_dia_kernel(int nx, int ny, int nz, double* val)
{
    const long long n = (long long)nx * ny * nz;
    const long long sxy = (long long)nx * ny;
    for (long long row = blockIdx.x * (long long)blockDim.x + threadIdx.x; row < n;
         row += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(row % nx), y = (int)((row / nx) % ny), z = (int)(row / sxy);
        val[0 * n + row] = z > 0 ? -1.0 : 0.0;
        val[1 * n + row] = y > 0 ? -1.0 : 0.0;
        val[2 * n + row] = x > 0 ? -1.0 : 0.0;
        val[3 * n + row] = 6.0;
        val[4 * n + row] = x < nx - 1 ? -1.0 : 0.0;
        val[5 * n + row] = y < ny - 1 ? -1.0 : 0.0;
        val[6 * n + row] = z < nz - 1 ? -1.0 : 0.0;
    }
}

}  // namespace

int poisson7_csr(int nx, int ny, int nz, int32_t* rowptr, int32_t* col, double* val, void* stream)
{
    hipLaunchKernelGGL(poisson_csr_kernel, dim3(4096), dim3(256), 0, (hipStream_t)stream, nx, ny, nz, rowptr, col,
                       val);
    PERKS_HIP_TRY(hipGetLastError());
    PERKS_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return PERKS_OK;
}

int poisson7_dia(int nx, int ny, int nz, double* val, void* stream)
{
    hipLaunchKernelGGL(poisson_dia_kernel, dim3(4096), dim3(256), 0, (hipStream_t)stream, nx, ny, nz, val);
    PERKS_HIP_TRY(hipGetLastError());
    PERKS_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return PERKS_OK;
}

}  // namespace perks

extern "C" {

void perks_cg_opts_init(perks_cg_opts* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->tol = 1e-10;
    o->inner_tol = 1e-5;
    o->relative = 1;
    o->max_iter = 10000;
    o->precision = PERKS_CG_F64;
    o->x0_is_zero = 0;
}

int perks_cg_solve_csr(const perks_csr* A, const double* d_b, double* d_x, const perks_cg_opts* opts, void* stream,
                       perks_cg_report* report)
{
    if (!A || !A->rowptr || !A->col || !A->val) return PERKS_ERR_INVALID_ARG;
    return perks::cg_solve(0, A, nullptr, d_b, d_x, opts, stream, report);
}

int perks_cg_solve_dia(const perks_dia* A, const double* d_b, double* d_x, const perks_cg_opts* opts, void* stream,
                       perks_cg_report* report)
{
    if (!A || !A->offsets || !A->val) return PERKS_ERR_INVALID_ARG;
    return perks::cg_solve(1, nullptr, A, d_b, d_x, opts, stream, report);
}

int perks_cg_solve_csr_f32(const perks_csr_f32* A, const float* d_b, float* d_x, const perks_cg_opts* opts,
                           void* stream, perks_cg_report* report)
{
    if (!A || !A->rowptr || !A->col || !A->val) return PERKS_ERR_INVALID_ARG;
    return perks::cg_solve_f32(0, A, nullptr, d_b, d_x, opts, stream, report);
}

int perks_cg_solve_dia_f32(const perks_dia_f32* A, const float* d_b, float* d_x, const perks_cg_opts* opts,
                           void* stream, perks_cg_report* report)
{
    if (!A || !A->offsets || !A->val) return PERKS_ERR_INVALID_ARG;
    return perks::cg_solve_f32(1, nullptr, A, d_b, d_x, opts, stream, report);
}

int64_t perks_poisson7_nnz(int nx, int ny, int nz)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) return PERKS_ERR_INVALID_ARG;
    const int64_t n = (int64_t)nx * ny * nz;
    return 7 * n - 2 * ((int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny);
}

int perks_poisson7_csr(int nx, int ny, int nz, int32_t* d_rowptr, int32_t* d_col, double* d_val, void* stream)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !d_rowptr || !d_col || !d_val) return PERKS_ERR_INVALID_ARG;
    if (perks_poisson7_nnz(nx, ny, nz) > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
    return perks::poisson7_csr(nx, ny, nz, d_rowptr, d_col, d_val, stream);
}

int perks_poisson7_dia(int nx, int ny, int nz, int32_t* offsets, double* d_val, void* stream)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !offsets || !d_val) return PERKS_ERR_INVALID_ARG;
    if ((int64_t)nx * ny * nz > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
    const int32_t sxy = nx * ny;
    const int32_t o[7] = {-sxy, -nx, -1, 0, 1, nx, sxy};
    for (int d = 0; d < 7; ++d) offsets[d] = o[d];
    return perks::poisson7_dia(nx, ny, nz, d_val, stream);
}

}  // extern "C"
