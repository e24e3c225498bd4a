// storage_fmt.h -- element access for the shape-generic stencil kernels: the
// storage type in HBM and the arithmetic type C it is widened to.  FP16/BF16
// fields hold raw 16-bit patterns and compute in FP32; every time step's
// result is rounded to the storage format (nearest even, or stochastically
// with the counter-hash bits of storage16.h).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "storage16.h"

namespace perks {
namespace fmt {

struct f16s {
    uint16_t v;
};
struct bf16s {
    uint16_t v;
};

template <typename S>
struct Fmt;
// Raw is the register form of a loaded storage word: a kernel that keeps
// loads in flight holds Raw values and widens each when it is first used, so
// the load's latency overlaps the work in between.
template <>
struct Fmt<double> {
    using C = double;
    using Raw = double;
    __device__ static Raw ld_raw(const double* p, size_t i) { return p[i]; }
    __device__ static C widen(Raw w) { return w; }
    __device__ static C ld(const double* p, size_t i) { return p[i]; }
    __device__ static void st(double* p, size_t i, C v) { p[i] = v; }
    __device__ static C q(C v) { return v; }
    __device__ static C qsr(C v, uint32_t, uint32_t) { return v; }
};
template <>
struct Fmt<float> {
    using C = float;
    using Raw = float;
    __device__ static Raw ld_raw(const float* p, size_t i) { return p[i]; }
    __device__ static C widen(Raw w) { return w; }
    __device__ static C ld(const float* p, size_t i) { return p[i]; }
    __device__ static void st(float* p, size_t i, C v) { p[i] = v; }
    __device__ static C q(C v) { return v; }
    __device__ static C qsr(C v, uint32_t, uint32_t) { return v; }
};
template <>
struct Fmt<f16s> {
    using C = float;
    using Raw = uint32_t;
    __device__ static Raw ld_raw(const f16s* p, size_t i) { return p[i].v; }
    __device__ static C widen(Raw w) { return __half2float(__ushort_as_half((uint16_t)w)); }
    __device__ static C ld(const f16s* p, size_t i) { return __half2float(__ushort_as_half(p[i].v)); }
    __device__ static void st(f16s* p, size_t i, C v) { p[i].v = __half_as_ushort(__float2half_rn(v)); }
    // The value after a round trip through storage (every step is stored).
    __device__ static C q(C v) { return __half2float(__float2half_rn(v)); }
    // Stochastic rounding of cell idx at the step with hash key `key`.
    __device__ static C qsr(C v, uint32_t idx, uint32_t key)
    {
        return Store16<PERKS_F16>::widen(narrow_as<PERKS_F16, true>(v, idx, key));
    }
};
template <>
struct Fmt<bf16s> {
    using C = float;
    using Raw = uint32_t;
    __device__ static Raw ld_raw(const bf16s* p, size_t i) { return p[i].v; }
    __device__ static C widen(Raw w) { return __uint_as_float(w << 16); }
    __device__ static C ld(const bf16s* p, size_t i) { return __uint_as_float((uint32_t)p[i].v << 16); }
    // Round to nearest even; NaN stays a quiet NaN.
    __device__ static uint16_t bits(C v) { return (uint16_t)Store16<PERKS_BF16>::narrow(v); }
    __device__ static void st(bf16s* p, size_t i, C v) { p[i].v = bits(v); }
    __device__ static C q(C v) { return __uint_as_float((uint32_t)bits(v) << 16); }
    __device__ static C qsr(C v, uint32_t idx, uint32_t key)
    {
        return Store16<PERKS_BF16>::widen(narrow_as<PERKS_BF16, true>(v, idx, key));
    }
};

}  // namespace fmt
}  // namespace perks
