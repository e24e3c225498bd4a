// storage16.h -- FP16 / BF16 storage with FP32 arithmetic.
//
// Fields are stored as raw 16-bit patterns; every stored value is the FP32
// result rounded to nearest-even by gfx950's conversions (v_cvt_f16_f32,
// v_cvt_pk_f16_f32, v_cvt_pk_bf16_f32).  The bfloat16 conversion equals the
// bit form rne_bf16_bits below on all 2^32 inputs (NaN kept quiet with its
// payload), which is how the CPU restatement states it.  Two neighbouring
// cells travel in one 32-bit word: the cell with the even x index in the low
// half.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "perks_amd.h"

namespace perks {

typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Round-to-nearest-even float -> bfloat16 bits, NaN quieted with its payload.
__host__ __device__ __forceinline__ uint32_t rne_bf16_bits(uint32_t u)
{
    return ((u & 0x7FFFFFFFu) > 0x7F800000u) ? ((u >> 16) | 0x40u) : ((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

template <int FMT>
struct Store16;

template <>
struct Store16<PERKS_F16> {
    __device__ static float widen(uint32_t h) { return __half2float(__ushort_as_half((unsigned short)h)); }
    __device__ static uint32_t narrow(float v) { return (uint32_t)__half_as_ushort(__float2half_rn(v)); }
    __device__ static uint32_t narrow2(float lo, float hi)
    {
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t{lo, hi}), f16x2_t));
    }
};

template <>
struct Store16<PERKS_BF16> {
    __device__ static float widen(uint32_t h) { return __uint_as_float(h << 16); }
    __device__ static uint32_t narrow(float v) { return narrow2(v, v) & 0xFFFFu; }
    __device__ static uint32_t narrow2(float lo, float hi)
    {
        return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t{lo, hi}), bf16x2_t));
    }
};

template <int FMT>
__device__ __forceinline__ float widen_lo(uint32_t w)
{
    return FMT == PERKS_BF16 ? __uint_as_float(w << 16) : Store16<FMT>::widen(w & 0xFFFFu);
}

template <int FMT>
__device__ __forceinline__ float widen_hi(uint32_t w)
{
    return FMT == PERKS_BF16 ? __uint_as_float(w & 0xFFFF0000u) : Store16<FMT>::widen(w >> 16);
}

template <int FMT>
__device__ __forceinline__ uint32_t pack2(float lo, float hi)
{
    return Store16<FMT>::narrow2(lo, hi);
}

// The value a result takes after a round trip through storage.
template <int FMT>
__device__ __forceinline__ float round16(float v)
{
    return Store16<FMT>::widen(Store16<FMT>::narrow(v));
}

// ---- stochastic rounding -------------------------------------------------------
// Random bits of cell `idx` (linear index mod 2^32) at time step t:
// lowbias32(idx ^ key(t)), key(t) = lowbias32(seed ^ t * 0x9E3779B9), with
// lowbias32 C. Wellons' public-domain integer hash (two odd multiplies and
// three xor-shifts).  The key is uniform across a launch's lanes.
__host__ __device__ __forceinline__ uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__host__ __device__ __forceinline__ uint32_t sr_key(uint32_t step, uint32_t seed)
{
    return lowbias32(seed ^ (step * 0x9E3779B9u));
}

// v rounded up or down to a neighbour in the storage format with probability
// proportional to the distance, using the random bits r.  BF16: r's low 16
// bits are added to the 16 bits that are dropped.  FP16: in the FP16 normal
// range r's low 13 bits are added to the 13 dropped bits (above 65504 the
// result is inf); subnormal results, zero, inf and NaN round to nearest even.
template <int FMT>
__device__ __forceinline__ uint32_t narrow_sr(float v, uint32_t r);

template <>
__device__ __forceinline__ uint32_t narrow_sr<PERKS_BF16>(float v, uint32_t r)
{
    const uint32_t u = __float_as_uint(v);
    return __builtin_isnan(v) ? ((u >> 16) | 0x40u) : ((u + (r & 0xFFFFu)) >> 16);
}

// FP16: an exponent field of 113..254 is the FP16 normal range and above.
// There the truncated sum t has 13 zero low bits, so converting it is exact,
// or gives +-inf once |t| > 65504 (the next such value, 65536, rounds to inf):
// the overflow rule above by the conversion itself.  Other inputs (FP16
// subnormal results, zero, inf, NaN) are converted from v, to nearest even.
template <>
__device__ __forceinline__ uint32_t narrow_sr<PERKS_F16>(float v, uint32_t r)
{
    const uint32_t u = __float_as_uint(v);
    const uint32_t e = (u >> 23) & 0xFFu;
    const uint32_t t = (u + (r & 0x1FFFu)) & 0xFFFFE000u;
    return Store16<PERKS_F16>::narrow(e - 113u <= 141u ? __uint_as_float(t) : v);
}

// Rounds v to the storage format: nearest even, or stochastically with the
// bits of (idx, key).
template <int FMT, bool SR>
__device__ __forceinline__ uint32_t narrow_as(float v, uint32_t idx, uint32_t key)
{
    if constexpr (SR) return narrow_sr<FMT>(v, lowbias32(idx ^ key));
    else return Store16<FMT>::narrow(v);
}

}  // namespace perks
