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
