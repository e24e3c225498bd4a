// fill.hip -- seeded uniform inputs generated directly in HBM.
//
// Element i is lo + (hi-lo) * u53(splitmix64(seed + (i+1)*golden)), the same
// counter-based sequence as the host generator, so device-generated and
// host-generated inputs are bit-identical (no FMA contraction: the multiply
// and the add round separately).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "perks_amd.h"

namespace perks {
namespace {

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long seed, unsigned long long i)
{
    unsigned long long z = seed + (i + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void __launch_bounds__(256) fill_uniform_kernel(T* __restrict__ d, size_t n, unsigned long long seed,
                                                           double lo, double w)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double u = (double)(splitmix64(seed, i) >> 11) * 0x1.0p-53;
        d[i] = (T)(lo + w * u);
    }
}

// 16-bit storage: the float value of the sequence, rounded to nearest even.
template <int BF16>
__global__ void __launch_bounds__(256) fill_uniform16_kernel(uint16_t* __restrict__ d, size_t n,
                                                             unsigned long long seed, double lo, double w)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double u = (double)(splitmix64(seed, i) >> 11) * 0x1.0p-53;
        const float f = (float)(lo + w * u);
        if (BF16) {
            const uint32_t b = __float_as_uint(f);
            d[i] = (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
        } else {
            d[i] = __half_as_ushort(__float2half_rn(f));
        }
    }
}

template <int BF16>
int fill_uniform16(uint16_t* d, size_t n, uint64_t seed, double lo, double hi, void* stream)
{
    if (n == 0) return PERKS_OK;
    if (!d) return PERKS_ERR_INVALID_ARG;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_uniform16_kernel<BF16>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, d, n,
                       (unsigned long long)seed, lo, hi - lo);
    if (hipGetLastError() != hipSuccess) return PERKS_ERR_HIP;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? PERKS_OK : PERKS_ERR_HIP;
}

template <typename T>
int fill_uniform(T* d, size_t n, uint64_t seed, double lo, double hi, void* stream)
{
    if (n == 0) return PERKS_OK;
    if (!d) return PERKS_ERR_INVALID_ARG;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_uniform_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, d, n,
                       (unsigned long long)seed, lo, hi - lo);
    if (hipGetLastError() != hipSuccess) return PERKS_ERR_HIP;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? PERKS_OK : PERKS_ERR_HIP;
}

}  // namespace
}  // namespace perks

extern "C" int perks_fill_uniform_f64(double* d, size_t n, uint64_t seed, double lo, double hi, void* stream)
{
    return perks::fill_uniform<double>(d, n, seed, lo, hi, stream);
}

extern "C" int perks_fill_uniform_f32(float* d, size_t n, uint64_t seed, double lo, double hi, void* stream)
{
    return perks::fill_uniform<float>(d, n, seed, lo, hi, stream);
}

extern "C" int perks_fill_uniform_f16(uint16_t* d, size_t n, uint64_t seed, double lo, double hi, void* stream)
{
    return perks::fill_uniform16<0>(d, n, seed, lo, hi, stream);
}

extern "C" int perks_fill_uniform_bf16(uint16_t* d, size_t n, uint64_t seed, double lo, double hi, void* stream)
{
    return perks::fill_uniform16<1>(d, n, seed, lo, hi, stream);
}
