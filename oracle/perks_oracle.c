/*
 * perks_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see perks_oracle.h).
 *
 * Build: oracle/Makefile, `-O3 -ffp-contract=off` with no FMA target flags so
 * every product and sum is rounded exactly once, in the order written.
 */
#include "perks_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Inputs                                                                    */
/* ------------------------------------------------------------------------ */

/* SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
 * generators", OOPSLA 2014; finaliser constants from Vigna's splitmix64.c),
 * used counter-style: the state for element i is seed + (i+1)*golden, so any
 * element can be generated independently (on the host or on the GPU). */
uint64_t po_splitmix64(uint64_t seed, uint64_t index)
{
    uint64_t z = seed + (index + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double u53(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

void po_fill_uniform_f64(double* out, size_t n, uint64_t seed, double lo, double hi)
{
    const double w = hi - lo;
    for (size_t i = 0; i < n; ++i) out[i] = lo + w * u53(po_splitmix64(seed, i));
}

void po_fill_uniform_f32(float* out, size_t n, uint64_t seed, double lo, double hi)
{
    const double w = hi - lo;
    for (size_t i = 0; i < n; ++i) out[i] = (float)(lo + w * u53(po_splitmix64(seed, i)));
}

/* stencil/3dstencil/common/common.hpp:24-27 (get_random) filled in z,y,x order
 * (common.hpp:59-69). */
void po_fill_libc_rand_f64(double* out, size_t n, unsigned seed)
{
    srand(seed);
    for (size_t i = 0; i < n; ++i) out[i] = (double)rand() / (double)(RAND_MAX - 1);
}

void po_fill_libc_rand_f32(float* out, size_t n, unsigned seed)
{
    srand(seed);
    for (size_t i = 0; i < n; ++i) out[i] = (float)rand() / (float)(RAND_MAX - 1);
}

/* ------------------------------------------------------------------------ */
/* Stencil step                                                              */
/* ------------------------------------------------------------------------ */

/* Neighbour index along one axis.  Clamp = replicate the edge, as the gold's
 * MIN/MAX index expressions do (3d7pt_gold_gen.cpp:21-26).  Zero = the cell
 * lies in a halo that holds 0 (returns -1). */
static inline int nb(int i, int d, int n, int bc)
{
    int j = i + d;
    if (j < 0) return bc == PO_BC_CLAMP ? 0 : -1;
    if (j >= n) return bc == PO_BC_CLAMP ? n - 1 : -1;
    return j;
}

#define DEFINE_STEP(SUF, T)                                                                   \
static void step_planes_##SUF(const T* in, T* out, int nz, int ny, int nx, const po_stencil* st, \
                              int zb, int ze)                                                 \
{                                                                                             \
    const size_t plane = (size_t)ny * nx;                                                     \
    const int bc = st->boundary;                                                              \
    for (int z = zb; z < ze; ++z)                                                             \
        for (int y = 0; y < ny; ++y)                                                          \
            for (int x = 0; x < nx; ++x) {                                                    \
                int in_dbl = 0;                                                               \
                T acc_t = 0;                                                                  \
                double acc_d = 0.0;                                                           \
                for (int g = 0; g < st->ngroups; ++g) {                                       \
                    const po_group* gr = &st->groups[g];                                      \
                    T s = 0;                                                                  \
                    for (int p = 0; p < gr->npts; ++p) {                                      \
                        int zz = nb(z, gr->off[p][0], nz, bc);                                \
                        int yy = nb(y, gr->off[p][1], ny, bc);                                \
                        int xx = nb(x, gr->off[p][2], nx, bc);                                \
                        T v = (zz < 0 || yy < 0 || xx < 0) ? (T)0                             \
                                                           : in[zz * plane + (size_t)yy * nx + xx]; \
                        s = (p == 0) ? v : (T)(s + v);                                        \
                    }                                                                         \
                    if (sizeof(T) == sizeof(double) || !gr->prod_f64) {                       \
                        T prod = (T)((T)gr->coef * s);                                        \
                        if (g == 0) acc_t = prod;                                             \
                        else if (in_dbl) acc_d = acc_d + (double)prod;                        \
                        else acc_t = (T)(acc_t + prod);                                       \
                    } else {                                                                  \
                        double prod = gr->coef * (double)s;                                   \
                        if (g == 0) acc_d = prod;                                             \
                        else if (in_dbl) acc_d = acc_d + prod;                                \
                        else acc_d = (double)acc_t + prod;                                    \
                        in_dbl = 1;                                                           \
                    }                                                                         \
                }                                                                             \
                T r;                                                                          \
                if (in_dbl) r = (T)(st->divisor != 0.0 ? acc_d / st->divisor : acc_d);        \
                else r = (st->divisor != 0.0) ? (T)(acc_t / (T)st->divisor) : acc_t;          \
                out[z * plane + (size_t)y * nx + x] = r;                                      \
            }                                                                                 \
}                                                                                             \
                                                                                              \
void po_stencil_step_##SUF(const T* in, T* out, int nz, int ny, int nx, const po_stencil* st) \
{                                                                                             \
    step_planes_##SUF(in, out, nz, ny, nx, st, 0, nz);                                        \
}                                                                                             \
                                                                                              \
static void iterate_##SUF(const T* in, T* out, int nz, int ny, int nx, const po_stencil* st,  \
                          int steps, int threads)                                             \
{                                                                                             \
    const size_t n = (size_t)nz * ny * nx;                                                    \
    if (steps <= 0) { memcpy(out, in, n * sizeof(T)); return; }                               \
    /* Ping-pong so that the last step writes `out` (j3d_gold_iterative's parity rule,      */\
    /* 3d7pt_gold_gen.cpp:43-64): odd step counts start in `out`, even ones in `tmp`.       */\
    T* tmp = (T*)malloc(n * sizeof(T));                                                       \
    T* dst = (steps % 2 == 1) ? out : tmp;                                                    \
    T* other = (steps % 2 == 1) ? tmp : out;                                                  \
    const T* src = in;                                                                        \
    for (int s = 0; s < steps; ++s) {                                                         \
        if (threads > 1) {                                                                    \
            _Pragma("omp parallel for num_threads(threads) schedule(static)")                \
            for (int z = 0; z < nz; ++z) step_planes_##SUF(src, dst, nz, ny, nx, st, z, z + 1); \
        } else {                                                                              \
            step_planes_##SUF(src, dst, nz, ny, nx, st, 0, nz);                               \
        }                                                                                     \
        src = dst;                                                                            \
        T* t = dst; dst = other; other = t;                                                   \
    }                                                                                         \
    free(tmp);                                                                                \
}                                                                                             \
void po_stencil_iterate_##SUF(const T* in, T* out, int nz, int ny, int nx, const po_stencil* st, int steps) \
{ iterate_##SUF(in, out, nz, ny, nx, st, steps, 1); }                                         \
void po_stencil_iterate_omp_##SUF(const T* in, T* out, int nz, int ny, int nx, const po_stencil* st, int steps, int threads) \
{ iterate_##SUF(in, out, nz, ny, nx, st, steps, threads); }                                   \
                                                                                              \
double po_rms_error_##SUF(const T* a, const T* b, int ny, int nx, int z0, int z1, int y0, int y1, \
                          int x0, int x1, double* max_err)                                    \
{                                                                                             \
    double err = 0.0, mx = 0.0;                                                               \
    for (int z = z0; z < z1; ++z)                                                             \
        for (int y = y0; y < y1; ++y)                                                         \
            for (int x = x0; x < x1; ++x) {                                                   \
                size_t i = ((size_t)z * ny + y) * nx + x;                                     \
                double e = (double)a[i] - (double)b[i];                                       \
                e = e < 0 ? -e : e;                                                           \
                err += e * e;                                                                 \
                if (e > mx) mx = e;                                                           \
            }                                                                                 \
    if (max_err) *max_err = mx;                                                               \
    double cnt = (double)(z1 - z0) * (double)(y1 - y0) * (double)(x1 - x0);                   \
    return cnt > 0 ? sqrt(err / cnt) : 0.0;                                                   \
}

DEFINE_STEP(f64, double)
DEFINE_STEP(f32, float)

/* Contracted update (this library's tune[6] = 1 option; no reference
 * counterpart): the same group order as step_planes_*, each group's sum of
 * points formed as there, the first product rounded on its own and every
 * later product fused into the running sum, acc = fma(coef, s, acc), one
 * rounding per term; then the optional division.  Groups that the gold
 * forms in double (prod_f64, FP32 only) are not supported: returns -1. */
#define DEFINE_FUSED(SUF, T, FMA)                                                             \
int po_stencil_iterate_fused_##SUF(const T* in, T* out, int nz, int ny, int nx, const po_stencil* st, int steps) \
{                                                                                             \
    for (int g = 0; g < st->ngroups; ++g)                                                     \
        if (sizeof(T) != sizeof(double) && st->groups[g].prod_f64) return -1;                 \
    const size_t plane = (size_t)ny * nx, n = (size_t)nz * plane;                             \
    if (steps <= 0) { memcpy(out, in, n * sizeof(T)); return 0; }                             \
    T* tmp = (T*)malloc(n * sizeof(T));                                                       \
    T* dst = (steps % 2 == 1) ? out : tmp;                                                    \
    T* other = (steps % 2 == 1) ? tmp : out;                                                  \
    const T* src = in;                                                                        \
    const int bc = st->boundary;                                                              \
    for (int it = 0; it < steps; ++it) {                                                      \
        for (int z = 0; z < nz; ++z)                                                          \
            for (int y = 0; y < ny; ++y)                                                      \
                for (int x = 0; x < nx; ++x) {                                                \
                    T acc = 0;                                                                \
                    for (int g = 0; g < st->ngroups; ++g) {                                   \
                        const po_group* gr = &st->groups[g];                                  \
                        T s = 0;                                                              \
                        for (int p = 0; p < gr->npts; ++p) {                                  \
                            int zz = nb(z, gr->off[p][0], nz, bc);                            \
                            int yy = nb(y, gr->off[p][1], ny, bc);                            \
                            int xx = nb(x, gr->off[p][2], nx, bc);                            \
                            T v = (zz < 0 || yy < 0 || xx < 0) ? (T)0                         \
                                                               : src[zz * plane + (size_t)yy * nx + xx]; \
                            s = (p == 0) ? v : (T)(s + v);                                    \
                        }                                                                     \
                        acc = (g == 0) ? (T)((T)gr->coef * s) : FMA((T)gr->coef, s, acc);     \
                    }                                                                         \
                    if (st->divisor != 0.0) acc = (T)(acc / (T)st->divisor);                  \
                    dst[z * plane + (size_t)y * nx + x] = acc;                                \
                }                                                                             \
        src = dst;                                                                            \
        T* t = dst; dst = other; other = t;                                                   \
    }                                                                                         \
    free(tmp);                                                                                \
    return 0;                                                                                 \
}

DEFINE_FUSED(f64, double, fma)
DEFINE_FUSED(f32, float, fmaf)

/* 7-point star, the gold's loop (3d7pt_gold_gen.cpp:9-40) with the edge
 * columns split off so the interior loop has no boundary tests.  Rows and
 * planes beyond a face are `zrow` (zeros) under the zero boundary and the
 * edge row or plane under clamp. */
static void star7_plane(const double* in, double* out, int nz, int ny, int nx, const double* k, int clamp,
                        const double* zrow, int z)
{
    const size_t plane = (size_t)ny * nx;
    const int zt = z + 1 < nz ? z + 1 : (clamp ? nz - 1 : -1);
    const int zb = z > 0 ? z - 1 : (clamp ? 0 : -1);
    for (int y = 0; y < ny; ++y) {
        const int yn = y + 1 < ny ? y + 1 : (clamp ? ny - 1 : -1);
        const int ys = y > 0 ? y - 1 : (clamp ? 0 : -1);
        const double* c = in + (size_t)z * plane + (size_t)y * nx;
        const double* N = yn < 0 ? zrow : in + (size_t)z * plane + (size_t)yn * nx;
        const double* S = ys < 0 ? zrow : in + (size_t)z * plane + (size_t)ys * nx;
        const double* T = zt < 0 ? zrow : in + (size_t)zt * plane + (size_t)y * nx;
        const double* B = zb < 0 ? zrow : in + (size_t)zb * plane + (size_t)y * nx;
        double* o = out + (size_t)z * plane + (size_t)y * nx;
        for (int side = 0; side < (nx > 1 ? 2 : 1); ++side) {  /* the two edge columns */
            const int x = side ? nx - 1 : 0;
            const double e = x + 1 < nx ? c[x + 1] : (clamp ? c[nx - 1] : 0.0);
            const double w = x > 0 ? c[x - 1] : (clamp ? c[0] : 0.0);
            double acc = k[0] * e;
            acc = acc + k[1] * w;
            acc = acc + k[2] * N[x];
            acc = acc + k[3] * S[x];
            acc = acc + k[4] * T[x];
            acc = acc + k[5] * B[x];
            acc = acc + k[6] * c[x];
            o[x] = acc;
        }
        for (int x = 1; x < nx - 1; ++x) {
            double acc = k[0] * c[x + 1];
            acc = acc + k[1] * c[x - 1];
            acc = acc + k[2] * N[x];
            acc = acc + k[3] * S[x];
            acc = acc + k[4] * T[x];
            acc = acc + k[5] * B[x];
            acc = acc + k[6] * c[x];
            o[x] = acc;
        }
    }
}

void po_star7_iterate_omp_f64(const double* in, double* out, int nz, int ny, int nx, const double* k, int boundary,
                              int steps, int threads)
{
    const size_t n = (size_t)nz * ny * nx;
    if (steps <= 0) { memcpy(out, in, n * sizeof(double)); return; }
    const int clamp = boundary == PO_BC_CLAMP;
    double* zrow = (double*)calloc((size_t)nx, sizeof(double));
    /* the parity rule of iterate_*: the last step writes `out` */
    double* tmp = (double*)malloc(n * sizeof(double));
    double* dst = (steps % 2 == 1) ? out : tmp;
    double* other = (steps % 2 == 1) ? tmp : out;
    const double* src = in;
    for (int s = 0; s < steps; ++s) {
        _Pragma("omp parallel for num_threads(threads) schedule(static)")
        for (int z = 0; z < nz; ++z) star7_plane(src, dst, nz, ny, nx, k, clamp, zrow, z);
        src = dst;
        double* t = dst; dst = other; other = t;
    }
    free(tmp);
    free(zrow);
}

/* ------------------------------------------------------------------------ */
/* Reduced-precision storage                                                 */
/* ------------------------------------------------------------------------ */

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* IEEE binary32 -> binary16, round to nearest even, with subnormals,
 * overflow to infinity and NaN kept quiet. */
uint16_t po_f32_to_f16(float f)
{
    const uint32_t u = f2u(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t absu = u & 0x7FFFFFFFu;
    if (absu >= 0x7F800000u) /* inf or nan */
        return (uint16_t)(sign | 0x7C00u | (absu > 0x7F800000u ? 0x0200u | ((absu >> 13) & 0x3FFu) : 0u));
    if (absu >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u); /* rounds to >= 65520: inf */
    if (absu < 0x38800000u) {
        /* result is subnormal or zero: value = absu_float / 2^-24 rounded */
        if (absu < 0x33000000u) return (uint16_t)sign; /* < 2^-25: rounds to 0 */
        const uint32_t e = absu >> 23;                 /* biased exponent */
        const uint32_t m = (absu & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - e;               /* 14 + (113 - e) */
        uint32_t r = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (r & 1u))) ++r;
        return (uint16_t)(sign | r);
    }
    /* normal: rebias exponent, round mantissa from 23 to 10 bits */
    uint32_t r = ((absu - 0x38000000u) >> 13);
    const uint32_t rem = absu & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
    return (uint16_t)(sign | r);
}

float po_f16_to_f32(uint16_t h)
{
    const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    const uint32_t e = ((uint32_t)h >> 10) & 0x1Fu;
    const uint32_t m = (uint32_t)h & 0x3FFu;
    if (e == 0) {
        if (m == 0) return u2f(sign);
        return (sign ? -1.0f : 1.0f) * ldexpf((float)m, -24);
    }
    if (e == 31) return u2f(sign | 0x7F800000u | (m << 13));
    return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

uint16_t po_f32_to_bf16(float f)
{
    const uint32_t u = f2u(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u); /* quiet nan */
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

float po_bf16_to_f32(uint16_t b) { return u2f((uint32_t)b << 16); }

static void iterate_u16(const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const po_stencil* st,
                        int steps, int bf16)
{
    const size_t n = (size_t)nz * ny * nx;
    float* a = (float*)malloc(n * sizeof(float));
    float* b = (float*)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) a[i] = bf16 ? po_bf16_to_f32(in[i]) : po_f16_to_f32(in[i]);
    for (int s = 0; s < steps; ++s) {
        po_stencil_step_f32(a, b, nz, ny, nx, st);
        for (size_t i = 0; i < n; ++i) {
            const uint16_t q = bf16 ? po_f32_to_bf16(b[i]) : po_f32_to_f16(b[i]);
            a[i] = bf16 ? po_bf16_to_f32(q) : po_f16_to_f32(q);
        }
    }
    for (size_t i = 0; i < n; ++i) out[i] = bf16 ? po_f32_to_bf16(a[i]) : po_f32_to_f16(a[i]);
    free(a);
    free(b);
}

void po_stencil_iterate_f16s(const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const po_stencil* st, int steps)
{
    iterate_u16(in, out, nz, ny, nx, st, steps, 0);
}

void po_stencil_iterate_bf16s(const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const po_stencil* st, int steps)
{
    iterate_u16(in, out, nz, ny, nx, st, steps, 1);
}

/* ---- stochastic rounding of 16-bit storage (see perks_oracle.h) ---------- */
uint32_t po_lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

uint32_t po_sr_bits(uint32_t idx, uint32_t step, uint32_t seed)
{
    return po_lowbias32(idx ^ po_lowbias32(seed ^ (step * 0x9E3779B9u)));
}

uint16_t po_f32_to_bf16_sr(float f, uint32_t r)
{
    const uint32_t u = f2u(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + (r & 0xFFFFu)) >> 16);
}

uint16_t po_f32_to_f16_sr(float f, uint32_t r)
{
    const uint32_t u = f2u(f);
    const uint32_t e = (u >> 23) & 0xFFu;
    if (e < 113u || e == 255u) return po_f32_to_f16(f);
    const uint32_t t = (u + (r & 0x1FFFu)) & 0xFFFFE000u;
    if ((t & 0x7FFFFFFFu) > 0x477FE000u) return (uint16_t)(((u >> 16) & 0x8000u) | 0x7C00u);
    return po_f32_to_f16(u2f(t)); /* exactly representable: conversion is exact */
}

void po_stencil_iterate_sr16(const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const po_stencil* st,
                             int steps, int fmt, uint32_t seed)
{
    const size_t n = (size_t)nz * ny * nx;
    float* a = (float*)malloc(n * sizeof(float));
    float* b = (float*)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) a[i] = fmt ? po_bf16_to_f32(in[i]) : po_f16_to_f32(in[i]);
    for (int s = 0; s < steps; ++s) {
        po_stencil_step_f32(a, b, nz, ny, nx, st);
        for (size_t i = 0; i < n; ++i) {
            const uint32_t r = po_sr_bits((uint32_t)i, (uint32_t)s, seed);
            const uint16_t q = fmt ? po_f32_to_bf16_sr(b[i], r) : po_f32_to_f16_sr(b[i], r);
            a[i] = fmt ? po_bf16_to_f32(q) : po_f16_to_f32(q);
        }
    }
    for (size_t i = 0; i < n; ++i) out[i] = fmt ? po_f32_to_bf16(a[i]) : po_f32_to_f16(a[i]);
    free(a);
    free(b);
}

/* ------------------------------------------------------------------------ */
/* Sparse matrices                                                           */
/* ------------------------------------------------------------------------ */

int64_t po_poisson7_nnz(int nx, int ny, int nz)
{
    int64_t n = (int64_t)nx * ny * nz;
    return 7 * n - 2 * ((int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny);
}

void po_poisson7_csr(int nx, int ny, int nz, int32_t* rowptr, int32_t* col, double* val)
{
    int64_t k = 0;
    const int64_t sxy = (int64_t)nx * ny;
    int64_t row = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x, ++row) {
                rowptr[row] = (int32_t)k;
                if (z > 0)      { col[k] = (int32_t)(row - sxy); val[k++] = -1.0; }
                if (y > 0)      { col[k] = (int32_t)(row - nx);  val[k++] = -1.0; }
                if (x > 0)      { col[k] = (int32_t)(row - 1);   val[k++] = -1.0; }
                col[k] = (int32_t)row; val[k++] = 6.0;
                if (x < nx - 1) { col[k] = (int32_t)(row + 1);   val[k++] = -1.0; }
                if (y < ny - 1) { col[k] = (int32_t)(row + nx);  val[k++] = -1.0; }
                if (z < nz - 1) { col[k] = (int32_t)(row + sxy); val[k++] = -1.0; }
            }
    rowptr[row] = (int32_t)k;
}

void po_poisson7_dia(int nx, int ny, int nz, int32_t* offsets, double* val)
{
    const int64_t n = (int64_t)nx * ny * nz;
    const int32_t sxy = nx * ny;
    const int32_t offs[7] = {-sxy, -nx, -1, 0, 1, nx, sxy};
    for (int d = 0; d < 7; ++d) offsets[d] = offs[d];
    int64_t row = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x, ++row) {
                val[0 * n + row] = z > 0 ? -1.0 : 0.0;
                val[1 * n + row] = y > 0 ? -1.0 : 0.0;
                val[2 * n + row] = x > 0 ? -1.0 : 0.0;
                val[3 * n + row] = 6.0;
                val[4 * n + row] = x < nx - 1 ? -1.0 : 0.0;
                val[5 * n + row] = y < ny - 1 ? -1.0 : 0.0;
                val[6 * n + row] = z < nz - 1 ? -1.0 : 0.0;
            }
}

/* y_i = sum over stored entries of row i, in storage order, starting from 0. */
/* CSR row sums in the library's fixed order (header: po_csr_spmv_f64).  The
 * same text serves double and float accumulation. */
#define PO_CSR_ROW(T, SUF)                                                                          \
    static T csr_row_##SUF(const int32_t* col, const T* val, int32_t k0, int32_t k1, const T* x)    \
    {                                                                                               \
        if (k1 - k0 <= PO_CSR_SHORT_ROW) {                                                          \
            T s = 0;                                                                                \
            for (int32_t k = k0; k < k1; ++k) s = s + val[k] * x[col[k]];                           \
            return s;                                                                               \
        }                                                                                           \
        T row = 0;                                                                                  \
        for (int32_t c0 = k0; c0 < k1; c0 += PO_CSR_ROW_CHUNK) {                                    \
            const int32_t c1 = k1 - c0 > PO_CSR_ROW_CHUNK ? c0 + PO_CSR_ROW_CHUNK : k1;             \
            T acc[64];                                                                              \
            for (int l = 0; l < 64; ++l) {                                                          \
                acc[l] = 0;                                                                         \
                for (int32_t k = c0 + l; k < c1; k += 64) acc[l] = acc[l] + val[k] * x[col[k]];     \
            }                                                                                       \
            for (int s = 32; s >= 1; s >>= 1)                                                       \
                for (int l = 0; l < s; ++l) acc[l] = acc[l] + acc[l + s];                           \
            row = c0 == k0 ? acc[0] : row + acc[0];                                                 \
        }                                                                                           \
        return row;                                                                                 \
    }
PO_CSR_ROW(double, f64)
PO_CSR_ROW(float, f32)

void po_csr_spmv_f64(int n, const int32_t* rowptr, const int32_t* col, const double* val, const double* x, double* y)
{
    for (int i = 0; i < n; ++i) y[i] = csr_row_f64(col, val, rowptr[i], rowptr[i + 1], x);
}

void po_csr_spmv_f32(int n, const int32_t* rowptr, const int32_t* col, const float* val, const float* x, float* y)
{
    for (int i = 0; i < n; ++i) y[i] = csr_row_f32(col, val, rowptr[i], rowptr[i + 1], x);
}

/* Diagonals in offset order; entries whose column falls outside [0,n) are skipped. */
void po_dia_spmv_f64(int n, int ndiag, const int32_t* offsets, const double* val, const double* x, double* y)
{
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int d = 0; d < ndiag; ++d) {
            int64_t j = (int64_t)i + offsets[d];
            if (j < 0 || j >= n) continue;
            s = s + val[(size_t)d * n + i] * x[j];
        }
        y[i] = s;
    }
}

/* Deterministic dot product.  The vector is cut into segments of 4096
 * elements.  Inside a segment, lane l (0..255) sums elements l, l+256, ...,
 * l+15*256 in that order starting from 0; the 256 lane sums are then folded
 * pairwise with strides 128, 64, ..., 1 (v[l] += v[l+stride]).  The segment
 * sums are combined the same way: lane l sums segments l, l+256, ... in order,
 * then the same fold.  The GPU reduction is built to this exact order, so the
 * FP64 CG residual history is reproducible bit for bit. */
#define RED_LANES 256
#define RED_PER_LANE 16
#define RED_SEG (RED_LANES * RED_PER_LANE)

static double fold256(double* v)
{
    for (int w = RED_LANES / 2; w >= 1; w >>= 1)
        for (int l = 0; l < w; ++l) v[l] = v[l] + v[l + w];
    return v[0];
}

static double reduce_partials(const double* p, size_t np)
{
    double v[RED_LANES];
    for (int l = 0; l < RED_LANES; ++l) {
        double acc = 0.0;
        for (size_t j = (size_t)l; j < np; j += RED_LANES) acc = acc + p[j];
        v[l] = acc;
    }
    return fold256(v);
}

double po_dot_f64(const double* a, const double* b, size_t n)
{
    size_t nseg = (n + RED_SEG - 1) / RED_SEG;
    if (nseg == 0) return 0.0;
    double* part = (double*)malloc(nseg * sizeof(double));
    for (size_t s = 0; s < nseg; ++s) {
        double v[RED_LANES];
        for (int l = 0; l < RED_LANES; ++l) {
            double acc = 0.0;
            for (int k = 0; k < RED_PER_LANE; ++k) {
                size_t i = s * RED_SEG + (size_t)k * RED_LANES + l;
                if (i < n) acc = acc + a[i] * b[i];
            }
            v[l] = acc;
        }
        part[s] = fold256(v);
    }
    double r = reduce_partials(part, nseg);
    free(part);
    return r;
}

/* Same order; float operands, exact double products, double sums. */
double po_dot_f32(const float* a, const float* b, size_t n)
{
    size_t nseg = (n + RED_SEG - 1) / RED_SEG;
    if (nseg == 0) return 0.0;
    double* part = (double*)malloc(nseg * sizeof(double));
    for (size_t s = 0; s < nseg; ++s) {
        double v[RED_LANES];
        for (int l = 0; l < RED_LANES; ++l) {
            double acc = 0.0;
            for (int k = 0; k < RED_PER_LANE; ++k) {
                size_t i = s * RED_SEG + (size_t)k * RED_LANES + l;
                if (i < n) acc = acc + (double)a[i] * (double)b[i];
            }
            v[l] = acc;
        }
        part[s] = fold256(v);
    }
    double r = reduce_partials(part, nseg);
    free(part);
    return r;
}

static void spmv(int fmt, int n, const int32_t* rowptr, const int32_t* col, const double* val, int ndiag,
                 const double* x, double* y)
{
    if (fmt == 0) po_csr_spmv_f64(n, rowptr, col, val, x, y);
    else po_dia_spmv_f64(n, ndiag, col, val, x, y);
}

/* conjugateGradient/cg_kernels.cuh:721-975 (order of operations):
 *   r = b + (-1)*A x ; r1 = r.r ; k = 1
 *   loop while !(r1 < thr) and k <= max_iter:
 *     k > 1 : beta = r1/r0 ; p = 1*r + beta*p      else p = r
 *     Ap = A p ; alpha = r1 / (p.Ap)
 *     x = alpha*p + x ; r = (-alpha)*Ap + r ; r0 = r1 ; r1 = r.r ; k++
 *   iterations = k-1 */
int po_cg_f64(int fmt, int n, const int32_t* rowptr, const int32_t* col, const double* val, int ndiag,
              const double* b, double* x, double tol, int relative, int max_iter,
              double* hist, int hist_cap)
{
    double* r = (double*)malloc((size_t)n * sizeof(double));
    double* p = (double*)malloc((size_t)n * sizeof(double));
    double* Ap = (double*)malloc((size_t)n * sizeof(double));
    spmv(fmt, n, rowptr, col, val, ndiag, x, Ap);
    for (int i = 0; i < n; ++i) r[i] = b[i] + (-1.0) * Ap[i];
    double thr = tol * tol;
    if (relative) thr = thr * po_dot_f64(b, b, (size_t)n);
    double r1 = po_dot_f64(r, r, (size_t)n), r0 = 0.0;
    if (hist && hist_cap > 0) hist[0] = r1;
    int k = 1;
    while (!(r1 < thr) && k <= max_iter) {
        if (k > 1) {
            double beta = r1 / r0;
            for (int i = 0; i < n; ++i) p[i] = 1.0 * r[i] + beta * p[i];
        } else {
            memcpy(p, r, (size_t)n * sizeof(double));
        }
        spmv(fmt, n, rowptr, col, val, ndiag, p, Ap);
        double pAp = po_dot_f64(p, Ap, (size_t)n);
        double alpha = r1 / pAp;
        double na = -alpha;
        for (int i = 0; i < n; ++i) x[i] = alpha * p[i] + x[i];
        for (int i = 0; i < n; ++i) r[i] = na * Ap[i] + r[i];
        r0 = r1;
        r1 = po_dot_f64(r, r, (size_t)n);
        if (hist && k < hist_cap) hist[k] = r1;
        ++k;
    }
    free(r); free(p); free(Ap);
    return k - 1;
}

/* ---- FP32-inner / FP64-refinement CG (see perks_oracle.h) ---------------- */
static void spmv_f32(int fmt, int n, const int32_t* rowptr, const int32_t* col, const float* val, int ndiag,
                     const float* x, float* y)
{
    for (int i = 0; i < n; ++i) {
        float s = 0.0f;
        if (fmt == 0) {
            s = csr_row_f32(col, val, rowptr[i], rowptr[i + 1], x);
        } else {
            for (int d = 0; d < ndiag; ++d) {
                int64_t j = (int64_t)i + col[d];
                if (j < 0 || j >= n) continue;
                s = s + val[(size_t)d * n + i] * x[j];
            }
        }
        y[i] = s;
    }
}

int po_cg_f32_refine(int fmt, int n, const int32_t* rowptr, const int32_t* col, const double* val, int ndiag,
                     const double* b, double* x, double tol, int relative, double inner_tol, int max_inner,
                     int max_outer, double* outer_hist, int hist_cap, int* outer_out)
{
    const size_t nval = fmt == 0 ? (size_t)rowptr[n] : (size_t)ndiag * n;
    float* valf = (float*)malloc(nval * sizeof(float));
    for (size_t i = 0; i < nval; ++i) valf[i] = (float)val[i];
    double* r = (double*)malloc((size_t)n * sizeof(double));
    double* Ax = (double*)malloc((size_t)n * sizeof(double));
    float* rf = (float*)malloc((size_t)n * sizeof(float));
    float* bf = (float*)malloc((size_t)n * sizeof(float));
    float* d = (float*)malloc((size_t)n * sizeof(float));
    float* p = (float*)malloc((size_t)n * sizeof(float));
    float* Ap = (float*)malloc((size_t)n * sizeof(float));
    const double bb = po_dot_f64(b, b, (size_t)n);
    const double thr = relative ? tol * tol * bb : tol * tol;
    double rr_prev = -1.0;
    int total = 0, outer = 0;
    for (outer = 0; outer <= max_outer; ++outer) {
        if (outer == 0) {
            memcpy(r, b, (size_t)n * sizeof(double)); /* x0 = 0 */
        } else {
            spmv(fmt, n, rowptr, col, val, ndiag, x, Ax);
            for (int i = 0; i < n; ++i) r[i] = b[i] + (-1.0) * Ax[i];
        }
        const double rr = po_dot_f64(r, r, (size_t)n);
        if (outer_hist && outer < hist_cap) outer_hist[outer] = rr;
        if (rr < thr || outer == max_outer || rr == rr_prev) break;
        rr_prev = rr;
        for (int i = 0; i < n; ++i) { bf[i] = (float)r[i]; rf[i] = bf[i]; d[i] = 0.0f; }
        double r1 = po_dot_f32(rf, rf, (size_t)n), r0 = 0.0;
        /* The inner solve stops at relative tolerance max(inner_tol, sqrt(thr / rr) / 2):
           never further than the outer target needs. */
        double tau2 = inner_tol * inner_tol;
        if (rr > 0.0 && 0.25 * thr / rr > tau2) tau2 = 0.25 * thr / rr;
        const double ithr = tau2 * po_dot_f32(bf, bf, (size_t)n);
        int k = 1;
        while (!(r1 < ithr) && k <= max_inner) {
            if (k > 1) {
                const float beta = (float)(r1 / r0);
                for (int i = 0; i < n; ++i) p[i] = rf[i] + beta * p[i];
            } else {
                memcpy(p, rf, (size_t)n * sizeof(float));
            }
            spmv_f32(fmt, n, rowptr, col, valf, ndiag, p, Ap);
            const double pAp = po_dot_f32(p, Ap, (size_t)n);
            const float alpha = (float)(r1 / pAp);
            const float na = -alpha;
            for (int i = 0; i < n; ++i) d[i] = alpha * p[i] + d[i];
            for (int i = 0; i < n; ++i) rf[i] = na * Ap[i] + rf[i];
            r0 = r1;
            r1 = po_dot_f32(rf, rf, (size_t)n);
            ++k;
        }
        total += k - 1;
        for (int i = 0; i < n; ++i) x[i] = x[i] + (double)d[i];
    }
    if (outer_out) *outer_out = outer;
    free(valf); free(r); free(Ax); free(rf); free(bf); free(d); free(p); free(Ap);
    return total;
}

/* ---- FP32 CG (see perks_oracle.h) ------------------------------------------ */
int po_cg_f32(int fmt, int n, const int32_t* rowptr, const int32_t* col, const float* val, int ndiag,
              const float* b, float* x, double thr2, int relative, int max_iter, int static_iter,
              double* hist, int hist_cap)
{
    float* r = (float*)malloc((size_t)n * sizeof(float));
    float* p = (float*)malloc((size_t)n * sizeof(float));
    float* Ap = (float*)malloc((size_t)n * sizeof(float));
    spmv_f32(fmt, n, rowptr, col, val, ndiag, x, Ap);
    for (int i = 0; i < n; ++i) r[i] = b[i] + (-1.0f) * Ap[i];
    const double thr = relative ? thr2 * po_dot_f32(b, b, (size_t)n) : thr2;
    double r1 = po_dot_f32(r, r, (size_t)n), r0 = 0.0;
    if (hist && hist_cap > 0) hist[0] = r1;
    int k = 1;
    while ((static_iter || !(r1 < thr)) && k <= max_iter) {
        if (k > 1) {
            const float beta = (float)(r1 / r0);
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        } else {
            memcpy(p, r, (size_t)n * sizeof(float));
        }
        spmv_f32(fmt, n, rowptr, col, val, ndiag, p, Ap);
        const double pAp = po_dot_f32(p, Ap, (size_t)n);
        const float alpha = (float)(r1 / pAp);
        const float na = -alpha;
        for (int i = 0; i < n; ++i) x[i] = alpha * p[i] + x[i];
        for (int i = 0; i < n; ++i) r[i] = na * Ap[i] + r[i];
        r0 = r1;
        r1 = po_dot_f32(r, r, (size_t)n);
        if (hist && k < hist_cap) hist[k] = r1;
        ++k;
    }
    free(r); free(p); free(Ap);
    return k - 1;
}
