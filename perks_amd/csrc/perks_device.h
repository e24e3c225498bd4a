// perks_device.h -- device-side helpers for gfx950 (CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace perks {

// DPP control codes (GFX9 encoding): whole-wave shift by one lane.
//   WAVE_SHL1: lane i receives lane i+1 (lane 63 keeps `old`)
//   WAVE_SHR1: lane i receives lane i-1 (lane 0 keeps `old`)
constexpr int kDppWaveShl1 = 0x130;
constexpr int kDppWaveShr1 = 0x138;

template <int CTRL>
__device__ __forceinline__ float dpp_move(float src, float old)
{
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_move(double src, double old)
{
    const long long s = __double_as_longlong(src);
    const long long o = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)s, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Whole-wave shift by one lane with zero fill (bound_ctrl): one DPP move per
// 32-bit half and no `old` operand to set up.  Used where the end lanes are
// ghost lanes whose values are never consumed.
template <int CTRL>
__device__ __forceinline__ float dpp_shift0(float src)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(src), CTRL, 0xF, 0xF, true));
}

template <int CTRL>
__device__ __forceinline__ double dpp_shift0(double src)
{
    const long long s = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_mov_dpp((int)s, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(s >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Value of `v` in lane `lane` (wave-uniform result).
__device__ __forceinline__ float read_lane(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// ---- raw buffer access ------------------------------------------------------
// A buffer descriptor carries a base address and a size in bytes; the hardware
// range-checks every access against it: a load past the end returns 0 and a
// store past the end is dropped.  The kernels use this instead of branches:
// lanes with nothing to store get an offset of kOob, and a plane outside the
// domain under the zero rule gets a descriptor of size 0, so its loads read 0.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr uint32_t kOob = 0x80000000u;
// Descriptor word 3 for gfx9-family targets (data format 32, no swizzle).
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kRsrcWord3);
}

// aux: cache-policy bits of the instruction (0 = default; 2 = nontemporal).
template <int AUX = 0>
__device__ __forceinline__ double buf_load(rsrc_t r, uint32_t off, double)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, AUX));
}

template <int AUX = 0>
__device__ __forceinline__ float buf_load(rsrc_t r, uint32_t off, float)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, AUX));
}

template <int AUX = 0>
__device__ __forceinline__ void buf_store(rsrc_t r, uint32_t off, double v)
{
    using v2i = int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, (int)off, 0, AUX);
}

template <int AUX = 0>
__device__ __forceinline__ void buf_store(rsrc_t r, uint32_t off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)off, 0, AUX);
}

// keep ? v : +0.0, without a branch.  Written as a bit mask on purpose: with a
// plain select the compiler may turn `cond ? 0 : load` into a conditional
// load followed by a full memory wait, which serialises the load stream.
__device__ __forceinline__ double keep_or_zero(double v, bool keep)
{
    const long long m = keep ? -1LL : 0LL;
    return __longlong_as_double(__double_as_longlong(v) & m);
}

__device__ __forceinline__ float keep_or_zero(float v, bool keep)
{
    const int m = keep ? -1 : 0;
    return __int_as_float(__float_as_int(v) & m);
}

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for every
// outstanding global load (s_waitcnt vmcnt(0)), which would drain the
// prefetch queue at each exchange; this waits for LDS operations only.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroup id remapped so that consecutive tiles share an XCD.  The
// dispatcher hands workgroup L to XCD L % 8; giving XCD k the contiguous tile
// range [k*G/8, (k+1)*G/8) keeps y/z neighbours (which re-read each other's
// halo rows) on one L2.  Identity when G is not a multiple of 8.
__device__ __forceinline__ int xcd_tile_id(int linear, int total)
{
    if ((total & 7) != 0) return linear;
    return (linear & 7) * (total >> 3) + (linear >> 3);
}

}  // namespace perks
