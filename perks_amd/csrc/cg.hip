// cg.hip -- persistent conjugate gradient (CSR and DIA; FP64, FP32, and FP32
// inner iterations inside an FP64 iterative-refinement loop).
//
// Replaces the reference's gpuConjugateGradient_cub (conjugateGradient/
// cg_kernels.cuh:620-986), its dispatcher (cg_dispatcher.cuh:559-916) and the
// merge-path SpMV it calls (cub/agent/agent_spmv_orig.cuh:666-1100,
// cub/device/dispatch/dispatch_spmv_orig.cuh:103, cub/agent/
// agent_segment_fixup.cuh:344): one cooperative launch runs the whole solve.
// The order of operations is the reference's (cg_kernels.cuh:721-975):
//     r = b + (-1)*A x ;  r1 = r.r
//     while !(r1 < thr):  p = 1*r + beta*p (beta = r1/r0 from the 2nd iteration)
//                         Ap = A p ; alpha = r1 / p.Ap
//                         x = alpha*p + x ; r = (-alpha)*Ap + r ; r0 = r1 ; r1 = r.r
//
// Work split.  Rows are grouped in segments of 4096.  The segments are cut
// into S = G * R slots (G workgroups, R = ceil(segments / G) rounds) by a
// merge-path split at segment granularity: the boundary cost of segment s is
// rowptr[4096 s] + 4096 s (entries plus rows), and slot c starts at the
// first segment whose cost reaches c/S of the total, so every slot holds
// about the same number of entries and rows whatever the row lengths (one
// segment per slot on a regular matrix).  Workgroup w runs slots w, w + G,
// w + 2G, ...: in each round the grid sweeps one contiguous band of the
// rows, so rows a plane apart are read while their owners are running (and,
// with the dispatcher dealing workgroups to the 8 XCDs round-robin, mostly
// through the same L2).  Inside a segment, lane t of the 256-thread
// workgroup owns rows t, t+256, ..., t+15*256.
//
// CSR rows by length (the order is fixed by the row length alone; the CPU
// oracle states it, po_csr_spmv_f64):
//   * up to 64 entries: the owning lane sums them in storage order;
//   * longer: the row is cut into chunks of up to 4096 entries and every
//     chunk is one work item for any wave of the grid (dealt round-robin, so
//     the skew of the row lengths spreads over the whole grid): lane l sums
//     entries l, l+64, ... of the chunk and the 64 lane sums are folded
//     pairwise (strides 32..1).  After a grid barrier the row's owner takes
//     chunk 0's value and adds the following chunks' in order (the fix-up),
//     and segments holding such rows get their dot product partials then.
//
// On-chip caching across iterations (the reference's cacheVector /
// cacheMatrix, cg_kernels.cuh:736-750,784-810,842-859,890-900,936-953): a
// workgroup always owns the same rows, so what only its own threads read can
// stay in LDS for the whole solve.  Per 256-row unit of the workgroup's rows,
// in this order while LDS lasts: Ap (never stored to HBM), x (written back
// at the end), p and r (still stored to HBM, where other workgroups' SpMV
// reads them, but never read back by the owner); then the matrix -- CSR: the
// column and value of the first entries of the workgroup's rows, DIA: the
// diagonals of its first rows -- filled by the first SpMV and read from LDS
// afterwards.  Every cached slot is written and read by one and the same
// thread, so caching needs no synchronisation and changes no arithmetic.
// The planner (plan_launch) sizes the caches: the LDS a workgroup may take
// with the grid's workgroups per CU, then the occupancy query decides, the
// counterpart of cg_dispatcher.cuh:688-748.
//
// Two grid barriers per iteration instead of the reference's five (three
// when long rows are present): p = r + beta*p is fused into the SpMV phase
// (a row's p is written by its owner; neighbours re-form the same value from
// r and the previous p, which is ping-ponged), and the x/r update is fused
// with the r.r partials.  Dot products are reduced in a fixed order instead
// of with atomics: lane l of a segment sums its 16 products in order, the 256
// lane sums are folded pairwise with strides 128..1, and every workgroup
// folds all segment sums the same way, so every workgroup holds the same
// scalar and the residual history is reproducible bit for bit (the oracle
// restates this order, po_dot_f64).
// Built with -ffp-contract=off: every multiply and add rounds on its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstring>
#include <mutex>
#include <vector>

#include "api_common.h"
#include "cg.h"
#include "grid_sync.h"

namespace perks {

namespace {

constexpr int kLanes = 256;
constexpr int kWaves = kLanes / 64;
constexpr int kPerLane = 16;
constexpr int kSeg = kLanes * kPerLane;
constexpr int kMaxDiag = 16;
constexpr int kCsrChunk = 8;    // entries of a short row loaded per predicated step
constexpr int kShortRow = 64;   // CSR rows up to this length: one lane
constexpr int kRowChunk = 4096; // longer rows: chunks of at most this many entries, one wave each
constexpr int kMaxWgPerCu = kCoopMaxWavesPerCu / 4;  // 256-thread workgroups

// Matrix formats the kernel is built for.  DIA7 is the 7-diagonal form with
// the diagonal count fixed at compile time (every load of a row issued at once).
enum { FMT_CSR = 0, FMT_DIA = 1, FMT_DIA7 = 2 };

// Vectors a workgroup may keep in LDS, in the order the planner fills them.
enum { VC_AP = 0, VC_X = 1, VC_P = 2, VC_R = 3, VC_N = 4 };

// Pointers into LDS carry its address space, so that a choice between a
// cached value and the one in HBM stays two loads of their own kinds (LDS and
// global) rather than one generic (flat) load through a selected pointer.
template <typename T>
using lds_ptr = __attribute__((address_space(3))) T*;

// The per-workgroup cache plan (the same for every workgroup of a launch).
struct CgCache {
    int units[VC_N];  // leading 256-row units of the workgroup's rows held in LDS, per vector
    int mat;          // CSR: leading entries of the workgroup's rows; DIA: leading rows
    int bytes;        // dynamic LDS per workgroup (slot table and caches)
    int table;        // of which the slot table
};

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
    // CSR rows longer than kShortRow ("split rows"), ascending: row j's chunks are
    // lchunk[j] .. lchunk[j+1]-1, chunk c belongs to row j = lcj[c], and
    // lpart[c] receives its value
    const uint8_t* seg_irr;  // CSR: 1 for a segment holding a row of more than kShortRow entries
    int balance_nnz;         // CSR with such segments: slots balance entries + rows (else rows)
    int nlong;
    int nchunk;
    const int32_t* lrow;
    const int32_t* lchunk;
    const int32_t* lcj;
    V* lpart;
    CgCache cache;
    int rounds;         // slots per workgroup (R)
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

// ---- reductions ----------------------------------------------------------
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

template <typename V>
__device__ __forceinline__ double prod_d(V a, V b)
{
    return (double)a * (double)b;  // exact for float operands; the product itself for double
}

// ---- the work split --------------------------------------------------------
// Merge-path cost at the start of segment s: entries before it plus rows
// before it when the matrix has segments of long rows, else rows (whole
// segments per slot: a split of near-equal segments by entries would give
// some slots two segments and others none).
template <typename V, int FMT>
__device__ __forceinline__ long long seg_cost(const CgArgs<V>& a, int s)
{
    const long long row = std::min((long long)s * kSeg, (long long)a.n);
    return FMT == FMT_CSR && a.balance_nnz ? (long long)a.rowptr[row] + row : row;
}

// First segment of slot c of S: the first whose cost reaches c/S of the total.
template <typename V, int FMT>
__device__ int slot_begin(const CgArgs<V>& a, int c, int S)
{
    if (c <= 0) return 0;
    if (c >= S) return a.nseg;
    const long long target = seg_cost<V, FMT>(a, a.nseg) * c / S;
    int lo = 0, hi = a.nseg;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (seg_cost<V, FMT>(a, mid) >= target) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Calls f(s, q) for the segments workgroup `me` owns, in its fixed order; q
// counts them (the index of the segment among the workgroup's own, which
// places its rows in the LDS caches).
template <typename V, class F>
__device__ __forceinline__ void for_owned(const CgArgs<V>& a, lds_ptr<const int> slots, const F& f)
{
    int q = 0;
    for (int k = 0; k < a.rounds; ++k) {
        const int s1 = slots[4 * k + 1];
        for (int s = slots[4 * k]; s < s1; ++s, ++q) f(s, q);
    }
}

// First long row at or after row `row` (index into lrow).
template <typename V>
__device__ int long_begin(const CgArgs<V>& a, long long row)
{
    int lo = 0, hi = a.nlong;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((long long)a.lrow[mid] >= row) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// ---- the workgroup's LDS caches ------------------------------------------
// The dynamic LDS starts with the workgroup's slot table, then the caches.
// Per round k: [4k] first and [4k+1] end segment of the workgroup's slot,
// [4k+2] and [4k+3] the range of its rows in the split-row list.
__host__ __device__ constexpr size_t slot_table_bytes(int rounds) { return ((size_t)rounds * 16 + 15) / 16 * 16; }

template <typename V>
struct Lds {
    lds_ptr<const int> slots;  // the slot table (slot_table_bytes)
    lds_ptr<V> vec[VC_N];
    int units[VC_N];
    int vmax;  // the largest of units[]: segments whose first unit is past it cache no vector
    lds_ptr<V> mval;
    lds_ptr<int32_t> mcol;
    int mat;
};

template <typename V, int FMT>
__device__ __forceinline__ Lds<V> lds_view(const CgArgs<V>& a, lds_ptr<unsigned char> dyn)
{
    Lds<V> c;
    c.slots = reinterpret_cast<lds_ptr<const int>>(dyn);
    size_t off = slot_table_bytes(a.rounds);
#pragma unroll
    for (int v = 0; v < VC_N; ++v) {
        c.vec[v] = reinterpret_cast<lds_ptr<V>>(dyn + off);
        c.units[v] = a.cache.units[v];
        off += (size_t)a.cache.units[v] * kLanes * sizeof(V);
    }
    c.vmax = 0;
#pragma unroll
    for (int v = 0; v < VC_N; ++v) c.vmax = c.units[v] > c.vmax ? c.units[v] : c.vmax;
    c.mat = a.cache.mat;
    c.mval = reinterpret_cast<lds_ptr<V>>(dyn + off);
    off += (size_t)c.mat * (FMT == FMT_CSR ? 1 : a.ndiag) * sizeof(V);
    c.mcol = reinterpret_cast<lds_ptr<int32_t>>(dyn + off);
    return c;
}

// Entry i of a vector for its owning thread (unit u of the workgroup's rows):
// from LDS when the unit is cached, else from HBM.
// Stores entry i of the row owned by thread `lane` (the split-row fix-up,
// where any thread of the workgroup assembles a row).
template <typename V>
__device__ __forceinline__ void vput_at(const Lds<V>& c, int v, int u, int lane, V* hbm, int i, V val, bool shared)
{
    if (u < c.units[v]) {
        c.vec[v][u * kLanes + lane] = val;
        if (shared) hbm[i] = val;
    } else {
        hbm[i] = val;
    }
}

// CACHED = false: the segment caches no vector, and the value is in HBM.
template <bool CACHED = true, typename V>
__device__ __forceinline__ V vget(const Lds<V>& c, int v, int u, const V* hbm, int i)
{
    if (CACHED && u < c.units[v]) return c.vec[v][u * kLanes + threadIdx.x];
    return hbm[i];
}
// Stores entry i: into LDS when cached (and to HBM too when other workgroups
// read the vector), else to HBM.
template <bool CACHED = true, typename V>
__device__ __forceinline__ void vput(const Lds<V>& c, int v, int u, V* hbm, int i, V val, bool shared)
{
    if (CACHED && u < c.units[v]) {
        c.vec[v][u * kLanes + threadIdx.x] = val;
        if (shared) hbm[i] = val;
    } else {
        hbm[i] = val;
    }
}

// ---- SpMV operands ----------------------------------------------------------
template <typename V>
struct OpVec {
    const V* v;
    __device__ __forceinline__ V operator()(int j) const { return v[j]; }
};
// p_j = r_j (first iteration) or r_j + beta * pold_j, formed on the fly.
template <typename V, bool FIRST>
struct OpP {
    const V* r;
    const V* pold;
    V beta;
    __device__ __forceinline__ V operator()(int j) const { return FIRST ? r[j] : r[j] + beta * pold[j]; }
};

// ---- CSR rows ----------------------------------------------------------------
// A short row [k0, k1) by its own lane, in storage order from zero.  All
// loads of a step are issued together (out-of-row slots read the row's first
// entry and the sum keeps its value through a select).  The entries come from
// `cols`/`vals` (HBM, or the workgroup's cached prefix in LDS); with fill,
// what is loaded is also stored into the prefix (entries from kb).
template <typename V, class Op, class CP, class VP>
__device__ __forceinline__ V csr_short_from(CP cols, VP vals, int off, int k0, int k1, const Lds<V>& c, int kb,
                                            bool fill, const Op& op)
{
    V s = (V)0;
    for (int k2 = k0; k2 < k1; k2 += kCsrChunk) {
        int cj[kCsrChunk];
        V cv[kCsrChunk];
#pragma unroll
        for (int u = 0; u < kCsrChunk; ++u) {
            const int k = (k2 + u < k1) ? k2 + u : k0;
            cj[u] = cols[k - off];
            cv[u] = vals[k - off];
        }
        if (fill) {
#pragma unroll
            for (int u = 0; u < kCsrChunk; ++u)
                if (k2 + u < k1) {
                    c.mcol[k2 + u - kb] = cj[u];
                    c.mval[k2 + u - kb] = cv[u];
                }
        }
#pragma unroll
        for (int u = 0; u < kCsrChunk; ++u) {
            const V prod = cv[u] * op(cj[u]);
            s = (k2 + u < k1) ? s + prod : s;
        }
    }
    return s;
}

// cached: the rows of every lane of the wave lie in the cached prefix.
template <typename V, class Op>
__device__ __forceinline__ V csr_short(const CgArgs<V>& a, const Lds<V>& c, int kb, int k0, int k1,
                                       bool cached, bool fill, const Op& op)
{
    if (cached && !fill) return csr_short_from<V>(c.mcol, c.mval, (int)kb, k0, k1, c, kb, false, op);
    return csr_short_from<V>(a.col, a.val, 0, k0, k1, c, kb, cached && fill, op);
}

// Entries [k0, k1) (at most kRowChunk) by a whole wave: lane l sums entries
// k0+l, k0+l+64, ... in order from zero; the 64 lane sums are folded
// pairwise with strides 32..1.  Every lane returns the result.  k0, k1,
// cached and fill are the same in every lane.
template <typename V, class Op>
__device__ __forceinline__ V csr_wave(const CgArgs<V>& a, const Lds<V>& c, int kb, int k0, int k1,
                                      bool cached, bool fill, const Op& op)
{
    constexpr int U = 4;  // entries per lane loaded together
    const int lane = threadIdx.x & 63;
    V acc = (V)0;
    for (int k2 = k0 + lane; k2 < k1; k2 += 64 * U) {
        int cj[U];
        V cv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = (k2 + 64 * u < k1) ? k2 + 64 * u : k2;
            if (cached && !fill) {
                cj[u] = c.mcol[k - kb];
                cv[u] = c.mval[k - kb];
            } else {
                cj[u] = a.col[k];
                cv[u] = a.val[k];
            }
        }
        if (cached && fill) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k2 + 64 * u < k1) {
                    c.mcol[k2 + 64 * u - kb] = cj[u];
                    c.mval[k2 + 64 * u - kb] = cv[u];
                }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const V prod = cv[u] * op(cj[u]);
            acc = (k2 + 64 * u < k1) ? acc + prod : acc;
        }
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) acc = acc + __shfl_down(acc, s, 64);
    return __shfl(acc, 0, 64);
}

// Row i of A times the operand, for the lane that owns row i, in a segment
// that may hold rows of more than kShortRow entries: those are left at 0 and
// flagged (their chunks are work items of any wave of the grid, and the
// row's value is assembled after a grid barrier).  Called by every lane of
// the wave.
template <typename V, class Op>
__device__ __forceinline__ V csr_row(const CgArgs<V>& a, const Lds<V>& c, int kb, int i, bool valid,
                                     bool fill, const Op& op, bool& lng)
{
    const int k0 = valid ? a.rowptr[i] : 0;
    const int k1 = valid ? a.rowptr[i + 1] : 0;
    lng = k1 - k0 > kShortRow;
    // The whole row lies in the cached prefix.  Rows use the cache only when
    // every lane's row does, so the choice of load path stays uniform over
    // the wave and its loads stay batched.
    const bool cached = (long long)k1 - kb <= (long long)c.mat;
    return csr_short(a, c, kb, k0, lng ? k0 : k1, __ballot(!cached) == 0, fill, op);
}

// ---- DIA rows -----------------------------------------------------------------
// Entries summed in diagonal order from zero; entries whose column falls
// outside [0, n) are skipped.  All loads are unconditional (out-of-range
// columns read row i itself and the sum keeps its value through a select).
// lr: the row's index among the workgroup's rows (cached when lr < mat).
// cached: every valid row of the wave lies in the cached rows (a choice
// uniform over the wave, as for CSR rows).
template <typename V, int FMT, class Op>
__device__ __forceinline__ V dia_row(const CgArgs<V>& a, const Lds<V>& c, int i, int lr, bool cached, bool fill,
                                     const Op& op)
{
    V s = (V)0;
    const int nd = FMT == FMT_DIA7 ? 7 : a.ndiag;
#pragma unroll
    for (int d = 0; d < (FMT == FMT_DIA7 ? 7 : kMaxDiag); ++d) {
        if (FMT != FMT_DIA7 && d >= nd) break;
        const long long j = (long long)i + a.offs[d];
        const bool in = (unsigned long long)j < (unsigned long long)a.n;
        const int jc = in ? (int)j : i;
        V mv;
        if (cached && !fill) mv = c.mval[d * c.mat + lr];
        else mv = a.val[(size_t)d * a.n + i];
        if (cached && fill) c.mval[d * c.mat + lr] = mv;
        const V prod = mv * op(jc);
        s = in ? s + prod : s;
    }
    return s;
}

// IRR: the segment may hold rows longer than kShortRow (else every row is a
// lane's, and no ballot separates one row's loads from the next row's).
// MC: the segment may lie in the workgroup's matrix cache (else its entries
// come straight from HBM, with no per-row choice either).
template <typename V, int FMT, bool IRR, bool MC, class Op>
__device__ __forceinline__ V row_value(const CgArgs<V>& a, const Lds<V>& c, int kb, int i, bool valid, int lr,
                                       bool fill, const Op& op, bool& lng)
{
    if constexpr (FMT == FMT_CSR && IRR) {
        return csr_row(a, c, kb, i, valid, MC && fill, op, lng);
    } else if constexpr (FMT == FMT_CSR) {
        lng = false;
        const int k0 = valid ? a.rowptr[i] : 0;
        const int k1 = valid ? a.rowptr[i + 1] : 0;
        if constexpr (!MC) return csr_short_from<V>(a.col, a.val, 0, k0, k1, c, kb, false, op);
        return csr_short(a, c, kb, k0, k1, __ballot((long long)k1 - kb > (long long)c.mat) == 0, fill, op);
    } else {
        lng = false;
        const bool cached = MC && __ballot(valid && lr >= c.mat) == 0;
        return valid ? dia_row<V, FMT>(a, c, i, lr, cached, MC && fill, op) : (V)0;
    }
}

// Whether segment s (the workgroup's q-th) may touch an LDS cache: a cached
// vector unit, or the matrix cache (CSR: entries from kb; DIA: leading rows).
template <typename V, int FMT>
__device__ __forceinline__ bool seg_cached(const CgArgs<V>& a, const Lds<V>& c, int s, int q, int kb)
{
    if (q * kPerLane < c.vmax) return true;
    if (c.mat == 0) return false;
    if constexpr (FMT == FMT_CSR) return (long long)a.rowptr[(long long)s * kSeg] - kb < (long long)c.mat;
    else return (long long)q * kSeg < (long long)c.mat;
}

// Every chunk of every long row, one wave per chunk across the grid.
template <typename V, class Op>
__device__ __forceinline__ void long_chunks(const CgArgs<V>& a, const Lds<V>& c, const Op& op)
{
    const int wave = threadIdx.x >> 6;
    for (int ci = blockIdx.x * kWaves + wave; ci < a.nchunk; ci += gridDim.x * kWaves) {
        const int j = a.lcj[ci];
        const int row = a.lrow[j];
        const int k0 = a.rowptr[row] + (ci - a.lchunk[j]) * kRowChunk;
        const int k1 = std::min(k0 + kRowChunk, (int)a.rowptr[row + 1]);
        const V v = csr_wave(a, c, 0, k0, k1, false, false, op);
        if ((threadIdx.x & 63) == 0) a.lpart[ci] = v;
    }
}

// The value of long row j: chunk 0's value plus each following chunk's, in order.
template <typename V>
__device__ __forceinline__ V long_value(const CgArgs<V>& a, int j)
{
    V v = a.lpart[a.lchunk[j]];
    for (int ci = a.lchunk[j] + 1; ci < a.lchunk[j + 1]; ++ci) v = v + a.lpart[ci];
    return v;
}

// ---- phases ---------------------------------------------------------------------
// r = b + (-1) * A x (r = b when x is zero); partials of r.r and b.b.  Also
// loads the cached units of x.  Segments holding long rows get their
// partials in fix_initial.
template <typename V, int FMT>
__device__ void phase_initial(const CgArgs<V>& a, const Lds<V>& c, int me, int kb, bool fill, double* lds)
{
    const int t = threadIdx.x;
    const OpVec<V> op{a.x};
    for_owned(a, c.slots, [&](int s, int q) {
        double rr = 0.0, bb = 0.0;
        int anyl = 0;
        auto rows = [&](auto irr) {
            for (int kk = 0; kk < kPerLane; ++kk) {
                const int i = s * kSeg + kk * kLanes + t;
                const int u = q * kPerLane + kk;
                const bool valid = i < a.n;
                bool lng = false;
                V ax = (V)0;
                if (!a.x_is_zero)
                    ax = row_value<V, FMT, decltype(irr)::value, true>(a, c, kb, i, valid, u * kLanes + t, fill, op,
                                                                        lng);
                if (valid) {
                    if (u < c.units[VC_X]) c.vec[VC_X][u * kLanes + t] = a.x[i];
                    if (!lng) {
                        const V bi = a.b[i];
                        const V ri = a.x_is_zero ? bi : bi + (V)(-1) * ax;
                        vput<true>(c, VC_R, u, a.r, i, ri, true);
                        rr = rr + prod_d(ri, ri);
                        bb = bb + prod_d(bi, bi);
                    }
                }
                anyl |= lng ? 1 : 0;
            }
        };
        if (FMT == FMT_CSR && a.seg_irr[s]) rows(std::true_type{});
        else rows(std::false_type{});
        const double prr = fold256(rr, lds);
        const double pbb = fold256(bb, lds);
        const int any = __syncthreads_or(anyl);
        if (t == 0 && !any) {
            a.part_a[s] = prr;
            a.part_b[s] = pbb;
        }
    });
    if constexpr (FMT == FMT_CSR)
        if (!a.x_is_zero) long_chunks(a, c, op);
}

// After a grid barrier: the long rows of this workgroup's segments (each by
// the thread that owns the row), then seg_fn(s, q) for every segment of the
// workgroup that holds one (a uniform walk of the sorted long-row list).
template <typename V, class Row, class Seg>
__device__ __forceinline__ void fix_long_rows(const CgArgs<V>& a, lds_ptr<const int> slots, const Row& row_fn,
                                              const Seg& seg_fn)
{
    // the split rows, spread over the workgroup's threads: row_fn(j, row, u,
    // lane) with u the row's unit and lane the thread that owns the row
    int q0 = 0;  // index of the slot's first segment among the workgroup's own
    for (int k = 0; k < a.rounds; ++k) {
        const int s0 = slots[4 * k], s1 = slots[4 * k + 1];
        for (int j = slots[4 * k + 2] + (int)threadIdx.x; j < slots[4 * k + 3]; j += kLanes) {
            const int row = a.lrow[j];
            const int s = row / kSeg;
            const int local = row - s * kSeg;
            row_fn(j, row, (q0 + s - s0) * kPerLane + (local >> 8), local & (kLanes - 1));
        }
        q0 += s1 - s0;
    }
    __syncthreads();  // cached values written for other threads' rows
    for_owned(a, slots, [&](int s, int q) {
        if (a.seg_irr[s]) seg_fn(s, q);
    });
}

template <typename V, int FMT>
__device__ void fix_initial(const CgArgs<V>& a, const Lds<V>& c, int me, double* lds)
{
    fix_long_rows(
        a, c.slots,
        [&](int j, int row, int u, int lane) {
            const V bi = a.b[row];
            const V ri = bi + (V)(-1) * long_value(a, j);
            vput_at(c, VC_R, u, lane, a.r, row, ri, true);
        },
        [&](int s, int q) {
            double rr = 0.0, bb = 0.0;
            for (int kk = 0; kk < kPerLane; ++kk) {
                const int i = s * kSeg + kk * kLanes + threadIdx.x;
                if (i < a.n) {
                    const V ri = vget(c, VC_R, q * kPerLane + kk, a.r, i);
                    const V bi = a.b[i];
                    rr = rr + prod_d(ri, ri);
                    bb = bb + prod_d(bi, bi);
                }
            }
            const double prr = fold256(rr, lds);
            const double pbb = fold256(bb, lds);
            if (threadIdx.x == 0) {
                a.part_a[s] = prr;
                a.part_b[s] = pbb;
            }
        });
}

// Phase 1 of an iteration: p = r + beta*p (stored into pnew), Ap = A p,
// partials of p.Ap into part_a.
template <typename V, int FMT, bool FIRST>
__device__ void phase_spmv(const CgArgs<V>& a, const Lds<V>& c, int me, int kb, const V* pold, V* pnew, V beta,
                           bool fill, double* lds)
{
    const int t = threadIdx.x;
    const OpP<V, FIRST> op{a.r, pold, beta};
    for_owned(a, c.slots, [&](int s, int q) {
        double acc = 0.0;
        int anyl = 0;
        auto rows = [&](auto irr, auto cached) {
            constexpr bool C = decltype(cached)::value;
#pragma unroll 2
            for (int kk = 0; kk < kPerLane; ++kk) {
                const int i = s * kSeg + kk * kLanes + t;
                const int u = q * kPerLane + kk;
                const bool valid = i < a.n;
                V pi = (V)0;
                if (valid) {
                    const V ri = vget<C>(c, VC_R, u, a.r, i);
                    pi = FIRST ? ri : ri + beta * vget<C>(c, VC_P, u, pold, i);
                }
                bool lng = false;
                const V api =
                    row_value<V, FMT, decltype(irr)::value, C>(a, c, kb, i, valid, u * kLanes + t, fill, op, lng);
                if (valid) {
                    vput<C>(c, VC_P, u, pnew, i, pi, true);
                    if (!lng) {
                        vput<C>(c, VC_AP, u, a.ap, i, api, false);
                        acc = acc + prod_d(pi, api);
                    }
                }
                anyl |= lng ? 1 : 0;
            }
        };
        const bool cv = seg_cached<V, FMT>(a, c, s, q, kb);
        if (FMT == FMT_CSR && a.seg_irr[s]) {
            if (cv) rows(std::true_type{}, std::true_type{});
            else rows(std::true_type{}, std::false_type{});
        } else {
            if (cv) rows(std::false_type{}, std::true_type{});
            else rows(std::false_type{}, std::false_type{});
        }
        const double ps = fold256(acc, lds);
        if constexpr (FMT == FMT_CSR) {
            const int any = __syncthreads_or(anyl);
            if (t == 0 && !any) a.part_a[s] = ps;
        } else {
            if (t == 0) a.part_a[s] = ps;
        }
    });
    if constexpr (FMT == FMT_CSR) long_chunks(a, c, op);
}

template <typename V, int FMT>
__device__ void fix_spmv(const CgArgs<V>& a, const Lds<V>& c, int me, const V* pnew, double* lds)
{
    fix_long_rows(
        a, c.slots,
        [&](int j, int row, int u, int lane) { vput_at(c, VC_AP, u, lane, a.ap, row, long_value(a, j), false); },
        [&](int s, int q) {
            double acc = 0.0;
            for (int kk = 0; kk < kPerLane; ++kk) {
                const int i = s * kSeg + kk * kLanes + threadIdx.x;
                if (i < a.n) {
                    const int u = q * kPerLane + kk;
                    acc = acc + prod_d(vget(c, VC_P, u, pnew, i), vget(c, VC_AP, u, a.ap, i));
                }
            }
            const double ps = fold256(acc, lds);
            if (threadIdx.x == 0) a.part_a[s] = ps;
        });
}

// Phase 2: x = alpha*p + x ; r = (-alpha)*Ap + r ; partials of r.r into part_b.
template <typename V, int FMT>
__device__ void phase_update(const CgArgs<V>& a, const Lds<V>& c, int me, const V* pnew, V alpha, double* lds)
{
    const int t = threadIdx.x;
    const V na = -alpha;
    for_owned(a, c.slots, [&](int s, int q) {
        double acc = 0.0;
        auto rows = [&](auto cached) {
            constexpr bool C = decltype(cached)::value;
#pragma unroll 4
            for (int kk = 0; kk < kPerLane; ++kk) {
                const int i = s * kSeg + kk * kLanes + t;
                if (i < a.n) {
                    const int u = q * kPerLane + kk;
                    const V pi = vget<C>(c, VC_P, u, pnew, i);
                    const V xi = alpha * pi + vget<C>(c, VC_X, u, a.x, i);
                    vput<C>(c, VC_X, u, a.x, i, xi, false);
                    const V ri = na * vget<C>(c, VC_AP, u, a.ap, i) + vget<C>(c, VC_R, u, a.r, i);
                    vput<C>(c, VC_R, u, a.r, i, ri, true);
                    acc = acc + prod_d(ri, ri);
                }
            }
        };
        if (q * kPerLane < c.vmax) rows(std::true_type{});
        else rows(std::false_type{});
        const double ps = fold256(acc, lds);
        if (t == 0) a.part_b[s] = ps;
    });
}

template <typename V, int FMT>
__global__ void __launch_bounds__(kLanes) cg_persistent_kernel(CgArgs<V> a)
{
    __shared__ double lds[kLanes];
    extern __shared__ __align__(16) unsigned char cg_dyn[];
    const int t = threadIdx.x;
    const int G = gridDim.x;
    const int me = blockIdx.x;
    const lds_ptr<unsigned char> dyn = (lds_ptr<unsigned char>)cg_dyn;
    const Lds<V> c = lds_view<V, FMT>(a, dyn);
    {
        // the slot table: round k's slot is me + k * G
        lds_ptr<int> tab = reinterpret_cast<lds_ptr<int>>(dyn);
        for (int k = t; k < a.rounds; k += kLanes) {
            const int s0 = slot_begin<V, FMT>(a, me + k * G, G * a.rounds);
            const int s1 = slot_begin<V, FMT>(a, me + k * G + 1, G * a.rounds);
            tab[4 * k] = s0;
            tab[4 * k + 1] = s1;
            tab[4 * k + 2] = a.nlong ? long_begin(a, (long long)s0 * kSeg) : 0;
            tab[4 * k + 3] = a.nlong ? long_begin(a, (long long)s1 * kSeg) : 0;
        }
        __syncthreads();
    }
    // the matrix cache holds the leading entries of the workgroup's first slot
    const int kb = FMT == FMT_CSR && a.rounds > 0 ? a.rowptr[std::min((long long)c.slots[0] * kSeg, (long long)a.n)]
                                                  : 0;
    const bool has_long = FMT == FMT_CSR && a.nlong > 0;

    // ---- r = b + (-1) * A x ; partials of r.r and b.b ----
    phase_initial<V, FMT>(a, c, me, kb, true, lds);
    bool fill = a.x_is_zero != 0;  // the first SpMV fills the matrix cache
    if (!grid_sync(a.sync, G, a.timeout)) return;
    if (has_long && !a.x_is_zero) {
        fix_initial<V, FMT>(a, c, me, lds);
        if (!grid_sync(a.sync, G, a.timeout)) return;
    }
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
        if (k == 1) phase_spmv<V, FMT, true>(a, c, me, kb, pold, pnew, (V)0, fill, lds);
        else phase_spmv<V, FMT, false>(a, c, me, kb, pold, pnew, (V)(r1 / r0), fill, lds);
        fill = false;
        if (!grid_sync(a.sync, G, a.timeout)) return;
        if (has_long) {
            fix_spmv<V, FMT>(a, c, me, pnew, lds);
            if (!grid_sync(a.sync, G, a.timeout)) return;
        }
        const double pap = reduce_parts(a.part_a, a.nseg, lds);
        const double alpha_d = r1 / pap;
        // ---- x = alpha*p + x ; r = (-alpha)*Ap + r ; partials of r.r ----
        phase_update<V, FMT>(a, c, me, pnew, (V)alpha_d, lds);
        if (!grid_sync(a.sync, G, a.timeout)) return;
        r0 = r1;
        r1 = reduce_parts(a.part_b, a.nseg, lds);
        if (me == 0 && t == 0 && a.hist && k < a.hist_cap) a.hist[k] = r1;
        ++k;
        V* tmp = pold;
        pold = pnew;
        pnew = tmp;
    }
    // ---- the cached units of x go back to HBM ----
    if (c.units[VC_X] > 0) {
        for_owned(a, c.slots, [&](int s, int q) {
            for (int kk = 0; kk < kPerLane; ++kk) {
                const int u = q * kPerLane + kk;
                const int i = s * kSeg + kk * kLanes + t;
                if (u < c.units[VC_X] && i < a.n) a.x[i] = c.vec[VC_X][u * kLanes + t];
            }
        });
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

// One wave per segment: flags the segments holding a row longer than
// kShortRow and counts those rows (they are split across the grid).
__global__ void classify_segments(const int32_t* __restrict__ rowptr, int n, int nseg, uint8_t* irr, int* nlong,
                                  int* nirr)
{
    const int lane = threadIdx.x & 63;
    for (int s = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); s < nseg; s += gridDim.x * (blockDim.x / 64)) {
        int longer = 0, lng = 0;
        for (int i = s * kSeg + lane; i < std::min(n, (s + 1) * kSeg); i += 64) {
            const int len = rowptr[i + 1] - rowptr[i];
            longer |= len > kShortRow ? 1 : 0;
            lng += len > kShortRow ? 1 : 0;
        }
        const unsigned long long any = __ballot(longer);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) lng += __shfl_down(lng, d, 64);
        if (lane == 0) {
            irr[s] = any ? 1 : 0;
            if (any) atomicAdd(nirr, 1);
            if (lng) atomicAdd(nlong, lng);
        }
    }
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
    double* small = nullptr;  // [0] rr, [1] bb ; int at [2] (iterations) and [4] (long rows)
    double* hist = nullptr;
    int hist_cap = 0;
    uint8_t* seg_irr = nullptr;  // CSR segment flags (nseg bytes)
    // long CSR rows: lrow, lchunk (nlong + 1), lcj and lpart (one double per chunk)
    int32_t* lrow = nullptr;
    int32_t* lchunk = nullptr;
    int32_t* lcj = nullptr;
    double* lpart = nullptr;
    size_t lrow_cap = 0, lchunk_cap = 0;
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

#define PERKS_TRY(expr)                              \
    do {                                             \
        const int perks_rc_ = (expr);                \
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
        PERKS_TRY(realloc_dev((void**)&s.seg_irr, (size_t)nseg));
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

struct LongRows {
    int nlong = 0;
    int nchunk = 0;
    int nirr = 0;  // segments holding rows of more than kShortRow entries
};

// The segment flags and the CSR rows longer than kShortRow with their chunk
// lists, on the device in S.  A classifying kernel runs first; the row
// offsets come to the host only when some row is that long.
int prepare_long_rows(const int32_t* d_rowptr, int n, CgScratch& S, hipStream_t stream, LongRows* out)
{
    *out = LongRows{};
    int* d_count = reinterpret_cast<int*>(S.small + 4);  // [0] long rows, [1] irregular segments
    const int nseg = (n + kSeg - 1) / kSeg;
    PERKS_HIP_TRY(hipMemsetAsync(d_count, 0, 2 * sizeof(int), stream));
    hipLaunchKernelGGL(classify_segments, dim3(std::min(4096, (nseg + 3) / 4)), dim3(256), 0, stream, d_rowptr, n,
                       nseg, S.seg_irr, d_count, d_count + 1);
    PERKS_HIP_TRY(hipGetLastError());
    int counts[2] = {0, 0};
    PERKS_HIP_TRY(hipMemcpyAsync(counts, d_count, sizeof(counts), hipMemcpyDeviceToHost, stream));
    PERKS_HIP_TRY(hipStreamSynchronize(stream));
    out->nirr = counts[1];
    const int count = counts[0];
    if (count == 0) return PERKS_OK;
    std::vector<int32_t> rp((size_t)n + 1);
    PERKS_HIP_TRY(hipMemcpy(rp.data(), d_rowptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost));
    std::vector<int32_t> lrow, lchunk{0}, lcj;
    for (int i = 0; i < n; ++i) {
        const int len = rp[(size_t)i + 1] - rp[i];
        if (len <= kShortRow) continue;
        const int j = (int)lrow.size();
        lrow.push_back(i);
        const int nch = (len + kRowChunk - 1) / kRowChunk;
        for (int c = 0; c < nch; ++c) lcj.push_back(j);
        lchunk.push_back(lchunk.back() + nch);
    }
    if (S.lrow_cap < lrow.size()) {
        S.lrow_cap = 0;
        PERKS_TRY(realloc_dev((void**)&S.lrow, sizeof(int32_t) * lrow.size()));
        PERKS_TRY(realloc_dev((void**)&S.lchunk, sizeof(int32_t) * (lrow.size() + 1)));
        S.lrow_cap = lrow.size();
    }
    if (S.lchunk_cap < lcj.size()) {
        S.lchunk_cap = 0;
        PERKS_TRY(realloc_dev((void**)&S.lcj, sizeof(int32_t) * lcj.size()));
        PERKS_TRY(realloc_dev((void**)&S.lpart, sizeof(double) * lcj.size()));
        S.lchunk_cap = lcj.size();
    }
    PERKS_HIP_TRY(hipMemcpy(S.lrow, lrow.data(), sizeof(int32_t) * lrow.size(), hipMemcpyHostToDevice));
    PERKS_HIP_TRY(hipMemcpy(S.lchunk, lchunk.data(), sizeof(int32_t) * lchunk.size(), hipMemcpyHostToDevice));
    PERKS_HIP_TRY(hipMemcpy(S.lcj, lcj.data(), sizeof(int32_t) * lcj.size(), hipMemcpyHostToDevice));
    out->nlong = (int)lrow.size();
    out->nchunk = (int)lcj.size();
    return PERKS_OK;
}

// ---- launch planning ---------------------------------------------------------
template <typename V, int FMT>
const void* kernel_ptr()
{
    return reinterpret_cast<const void*>(&cg_persistent_kernel<V, FMT>);
}

// The grid and the workgroups' LDS caches.  The grid is CUs x occupancy (at
// most kMaxWgPerCu per CU, at most one workgroup per segment).  A workgroup
// may take its CU's LDS divided by the grid's workgroups per CU, less the
// kernel's static LDS.  The budget goes to the vectors first (Ap, x, p, r:
// 256-row units up to the workgroup's average share of rows), then to the
// matrix; the occupancy query with that dynamic LDS must still admit the
// grid's workgroups per CU, else the budget shrinks and the plan is redone
// (cg_dispatcher.cuh:688-748 sizes the reference's caches the same way).
// LDS is allocated per workgroup in whole granules (static and dynamic
// together), so a share of exactly lds_per_cu / wpc can round up past the
// CU's LDS: then only wpc - 1 workgroups fit per CU, part of the cooperative
// grid never starts, and the first grid barrier waits for it until its time
// limit (at 256^3 and three workgroups per CU: 54384 B per workgroup, 54784
// in 512-byte granules, three of them 164352 B > 163840).  The share is
// rounded down to a multiple of every plausible granule (512 B, 1 KiB,
// 1280 B, 2 KiB), so wpc workgroups fit whichever the hardware uses.
static long long lds_share(long long per_wg)
{
    const long long a = per_wg / 2048 * 2048, b = per_wg / 1280 * 1280;
    return a < b ? a : b;
}

template <typename V, int FMT>
int plan_launch(const CgArgs<V>& a, const perks_cg_opts* o, int* grid, CgCache* cache)
{
    const void* fn = kernel_ptr<V, FMT>();
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    int occ = 0;
    PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kLanes, 0));
    if (occ <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
    if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
    // At most half the wave slots: 4 workgroups of 4 waves per CU keep enough
    // loads in flight, and the cooperative grid never needs every slot free.
    if (occ > kMaxWgPerCu) occ = kMaxWgPerCu;
    int G = caps.cu_count * occ;
    if (G > a.nseg) G = a.nseg > 0 ? a.nseg : 1;
    const int wpc = (G + caps.cu_count - 1) / caps.cu_count;
    hipFuncAttributes fa;
    PERKS_HIP_TRY(hipFuncGetAttributes(&fa, fn));
    const int rounds = (a.nseg + G - 1) / G;
    const long long table = (long long)slot_table_bytes(rounds);
    long long budget = lds_share(std::min<long long>(caps.lds_per_block, caps.lds_per_cu / wpc)) -
                       (long long)fa.sharedSizeBytes - table;
    budget = budget / 256 * 256;
    const long long unit = (long long)kLanes * sizeof(V);
    const long long share = (long long)((a.nseg + G - 1) / G) * kPerLane;  // 256-row units of a workgroup
    CgCache c{};
    for (;;) {
        std::memset(&c, 0, sizeof(c));
        long long left = budget > 0 ? budget : 0;
        if (o->cache_vector >= 0)
            for (int v = 0; v < VC_N; ++v) {
                c.units[v] = (int)std::min(share, left / unit);
                left -= c.units[v] * unit;
            }
        if (o->cache_matrix >= 0 && left > 0) {
            if (FMT == FMT_CSR) c.mat = (int)std::min<long long>(left / (long long)(sizeof(V) + 4), 0x7fffffffLL);
            else c.mat = (int)std::min(share * kLanes, left / ((long long)a.ndiag * (long long)sizeof(V)));
        }
        long long bytes = 0;
        for (int v = 0; v < VC_N; ++v) bytes += c.units[v] * unit;
        bytes += (long long)c.mat * (FMT == FMT_CSR ? (long long)sizeof(V) + 4 : (long long)a.ndiag * sizeof(V));
        c.bytes = (int)((bytes + 15) / 16 * 16);
        if (c.bytes == 0) break;
        c.bytes += (int)table;
        int occ2 = 0;
        if (c.bytes > 65536 &&
            hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, c.bytes) != hipSuccess) {
            (void)hipGetLastError();
            budget = (lds_share(65536) - (long long)fa.sharedSizeBytes - table) / 256 * 256;
            continue;
        }
        PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ2, fn, kLanes, (size_t)c.bytes));
        if (occ2 >= wpc) break;
        budget -= 4096;
    }
    if (c.bytes == 0) c.bytes = (int)table;
    c.table = (int)table;
    *grid = G;
    *cache = c;
    return PERKS_OK;
}

template <typename V, int FMT>
int launch_cg(CgArgs<V>& a, GridSync* gs, hipStream_t stream, const perks_cg_opts* o, int* grid_out,
              CgCache* cache_out)
{
    int G = 0;
    CgCache c;
    const int rc = plan_launch<V, FMT>(a, o, &G, &c);
    if (rc != PERKS_OK) return rc;
    a.cache = c;
    a.rounds = (a.nseg + G - 1) / G;
    a.sync = gs;
    a.timeout = grid_sync_timeout();
    void* args[] = {&a};
    PERKS_HIP_TRY(launch_persistent(kernel_ptr<V, FMT>(), dim3(G), dim3(kLanes), args, (size_t)c.bytes, stream));
    if (grid_out) *grid_out = G;
    if (cache_out) *cache_out = c;
    return PERKS_OK;
}

template <typename V>
int launch_cg_fmt(int fmt, CgArgs<V>& a, GridSync* gs, hipStream_t stream, const perks_cg_opts* o, int* grid_out,
                  CgCache* cache_out)
{
    if (fmt == FMT_CSR) return launch_cg<V, FMT_CSR>(a, gs, stream, o, grid_out, cache_out);
    if (a.ndiag == 7) return launch_cg<V, FMT_DIA7>(a, gs, stream, o, grid_out, cache_out);
    return launch_cg<V, FMT_DIA>(a, gs, stream, o, grid_out, cache_out);
}

void report_cache(perks_cg_report* rep, const CgCache& c, const LongRows& lr)
{
    rep->lds_bytes = c.bytes - c.table;
    rep->cache_units = c.units[VC_AP] + c.units[VC_X] + c.units[VC_P] + c.units[VC_R];
    rep->cache_matrix = c.mat;
    rep->long_rows = lr.nlong;
}

template <typename V>
void set_long(CgArgs<V>& a, const CgScratch& S, const LongRows& lr)
{
    a.seg_irr = S.seg_irr;
    a.balance_nnz = lr.nirr > 0 ? 1 : 0;
    a.nlong = lr.nlong;
    a.nchunk = lr.nchunk;
    a.lrow = S.lrow;
    a.lchunk = S.lchunk;
    a.lcj = S.lcj;
    a.lpart = reinterpret_cast<V*>(S.lpart);
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
    LongRows lr;
    if (fmt == FMT_CSR) {
        rc = prepare_long_rows(csr->rowptr, n, S, stream, &lr);
        if (rc != PERKS_OK) return rc;
    }

    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    int grid = 0;
    CgCache cache{};
    std::vector<double> hist_host;
    int total_iters = 0, outer = 0;
    double rr = 0.0, bb = 0.0;

    auto fill64 = [&](CgArgs<double>& a) {
        std::memset(&a, 0, sizeof(a));
        a.n = n;
        a.nseg = (n + kSeg - 1) / kSeg;  // the scratch may be sized for a larger problem
        if (fmt == FMT_CSR) {
            a.rowptr = csr->rowptr;
            a.col = csr->col;
            a.val = csr->val;
        } else {
            a.val = dia->val;
            a.ndiag = dia->ndiag;
            for (int d = 0; d < dia->ndiag; ++d) a.offs[d] = dia->offsets[d];
        }
        set_long(a, S, lr);
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
        rc = launch_cg_fmt<double>(fmt, a, lease.gs, stream, o, &grid, &cache);
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
            rc = launch_cg_fmt<double>(fmt, a, lease.gs, stream, o, &grid, nullptr);
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
            f.nseg = (n + kSeg - 1) / kSeg;  // the scratch may be sized for a larger problem
            if (fmt == FMT_CSR) {
                f.rowptr = csr->rowptr;
                f.col = csr->col;
            } else {
                f.ndiag = dia->ndiag;
                for (int d = 0; d < dia->ndiag; ++d) f.offs[d] = dia->offsets[d];
            }
            set_long(f, S, lr);
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
            rc = launch_cg_fmt<float>(fmt, f, lease.gs, stream, o, &grid, &cache);
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
    report_cache(rep, cache, lr);
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
    LongRows lr;
    if (fmt == FMT_CSR) {
        rc = prepare_long_rows(csr->rowptr, n, S, stream, &lr);
        if (rc != PERKS_OK) return rc;
    }
    CgArgs<float> f;
    std::memset(&f, 0, sizeof(f));
    f.n = n;
    f.nseg = (n + kSeg - 1) / kSeg;  // the scratch may be sized for a larger problem
    if (fmt == FMT_CSR) {
        f.rowptr = csr->rowptr;
        f.col = csr->col;
        f.val = csr->val;
    } else {
        f.ndiag = dia->ndiag;
        for (int d = 0; d < dia->ndiag; ++d) f.offs[d] = dia->offsets[d];
        f.val = dia->val;
    }
    set_long(f, S, lr);
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
    CgCache cache{};
    rc = launch_cg_fmt<float>(fmt, f, lease.gs, stream, o, &grid, &cache);
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
    report_cache(rep, cache, lr);
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
                    (void*)s.small, (void*)s.hist, (void*)s.seg_irr, (void*)s.lrow, (void*)s.lchunk, (void*)s.lcj, (void*)s.lpart})
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
