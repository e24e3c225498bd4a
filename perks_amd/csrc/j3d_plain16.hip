// j3d_plain16.hip -- per-step 3D 7-point Jacobi on FP16 / BF16 storage.
//
// Same z-march as the FP64/FP32 per-step kernel (j3d_plain.hip), but each lane
// owns two neighbouring x cells, loaded and stored as one 32-bit word, so a
// wave covers 128 columns and moves 256 B per row access.  Arithmetic is FP32
// in the reference gold's order; every result is rounded to the storage format
// (round to nearest even), as a stored field is after every time step.
//
// x neighbours: the west cell of the even cell is the odd cell of lane-1 (DPP
// shift), the east cell of the odd cell is the even cell of lane+1; inside the
// pair they are each other.  The two tile-edge columns (x0-1, x0+128) come
// from one extra 16-bit load per plane whose lanes r and 32+r fetch row r's
// west and east halo.  Needs an even row width (4-byte aligned pairs).
#include "j3d_kernels.h"
#include "storage16.h"

namespace perks {

namespace {

template <int FMT, int BC, int NY, bool HAS_DIV, bool SR>
struct Star1Plain16 {
    const uint16_t* __restrict__ in;
    uint16_t* __restrict__ out;
    int nz;
    size_t plane;            // elements per plane
    uint32_t plane_bytes;
    uint32_t roff[NY + 2];   // byte offsets of this lane's pair in rows y0-1 .. y0+NY (kOob: zero)
    uint32_t soff[NY];       // byte offsets of this lane's output pair (kOob: not stored)
    uint32_t hoff;           // this lane's halo element (row hr, west or east column)
    bool xfirst, xlast;      // this lane's pair holds x = 0 / x = nx-1 (clamp rule)
    uint32_t cidx[NY];       // stochastic rounding: linear index of the pair's even cell in plane 0
    uint32_t plane32;        // cells per plane, mod 2^32
    uint32_t key;            // stochastic rounding: this time step's hash key
    float k[8];

    // Issues the loads of plane zz; never waits.  Outside [0, nz) the clamp
    // rule reads the edge plane and the zero rule a size-0 descriptor.
    __device__ __forceinline__ void load(int zz, uint32_t (&v)[NY + 2], uint32_t& h) const
    {
        uint32_t bytes = plane_bytes;
        int zc = zz;
        if (zz < 0 || zz >= nz) {
            zc = zz < 0 ? 0 : nz - 1;
            if (BC == PERKS_BC_ZERO) bytes = 0;
        }
        const rsrc_t rs = make_rsrc(in + (size_t)zc * plane, bytes);
#pragma unroll
        for (int r = 0; r < NY + 2; ++r) v[r] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)roff[r], 0, 0);
        h = (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rs, (int)hoff, 0, 0);
    }

    __device__ __forceinline__ void compute(int z, const uint32_t (&bel)[NY + 2], const uint32_t (&cur)[NY + 2],
                                            uint32_t hcur, const uint32_t (&up)[NY + 2]) const
    {
        const rsrc_t rs = make_rsrc(out + (size_t)z * plane, plane_bytes);
        const float hf = Store16<FMT>::widen(hcur);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const float hw = read_lane(hf, r);
            const float he = read_lane(hf, 32 + r);
            const float c0 = widen_lo<FMT>(cur[r + 1]);
            const float c1 = widen_hi<FMT>(cur[r + 1]);
            float w0 = dpp_move<kDppWaveShr1>(c1, hw);  // odd cell of lane-1 (lane 0: west halo)
            float e1 = dpp_move<kDppWaveShl1>(c0, he);  // even cell of lane+1 (lane 63: east halo)
            if (BC == PERKS_BC_CLAMP) {
                w0 = xfirst ? c0 : w0;
                e1 = xlast ? c1 : e1;
            } else {
                w0 = xfirst ? 0.0f : w0;
                e1 = xlast ? 0.0f : e1;
            }
            const float v0 = star1_update<float, HAS_DIV>(k, c1, w0, widen_lo<FMT>(cur[r + 2]), widen_lo<FMT>(cur[r]),
                                                          widen_lo<FMT>(up[r + 1]), widen_lo<FMT>(bel[r + 1]), c0);
            const float v1 = star1_update<float, HAS_DIV>(k, e1, c0, widen_hi<FMT>(cur[r + 2]), widen_hi<FMT>(cur[r]),
                                                          widen_hi<FMT>(up[r + 1]), widen_hi<FMT>(bel[r + 1]), c1);
            const uint32_t i0 = (uint32_t)z * plane32 + cidx[r];
            const uint32_t packed = narrow_as<FMT, SR>(v0, i0, key) | (narrow_as<FMT, SR>(v1, i0 + 1u, key) << 16);
            __builtin_amdgcn_raw_buffer_store_b32((int)packed, rs, (int)soff[r], 0, 2);
        }
    }

    __device__ __forceinline__ void stage(int z, const uint32_t (&bel)[NY + 2], const uint32_t (&cur)[NY + 2],
                                          uint32_t hcur, const uint32_t (&up)[NY + 2], uint32_t (&nxt)[NY + 2],
                                          uint32_t& hnxt) const
    {
        load(z + 2, nxt, hnxt);
        compute(z, bel, cur, hcur, up);
    }
};

template <int FMT, int BC, int NY, bool HAS_DIV, bool SR>
__global__ void __launch_bounds__(256)
j3d_star1_plain16_kernel(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int nz, int ny, int nx,
                         int zchunk, int tiles_x, int tiles_y, KCoefs cf, uint32_t key)
{
    Star1Plain16<FMT, BC, NY, HAS_DIV, SR> S;
    load_coefs7<float>(cf, S.k);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tile = xcd_tile_id(blockIdx.x, gridDim.x);
    const int bx = tile % tiles_x;
    const int by = (tile / tiles_x) % tiles_y;
    const int bz = tile / (tiles_x * tiles_y);

    const int x0 = bx * 128;
    const int y0 = (by * 4 + wave) * NY;
    const int zb = bz * zchunk;
    const int ze = min(zb + zchunk, nz);
    S.in = in;
    S.out = out;
    S.nz = nz;
    S.plane = (size_t)ny * nx;
    S.plane_bytes = (uint32_t)(S.plane * 2);
    S.plane32 = (uint32_t)S.plane;
    S.key = key;
    const int x = x0 + 2 * lane;  // even cell of this lane's pair
    const bool xin = x < nx;      // nx is even: the pair is wholly in or out
    S.xfirst = x == 0;
    S.xlast = x + 1 == nx - 1;
#pragma unroll
    for (int r = 0; r < NY + 2; ++r) {
        const int yy = y0 + r - 1;
        const bool yin = yy >= 0 && yy < ny;
        const int yc = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
        const uint32_t off = (uint32_t)((yc * nx + (xin ? x : nx - 2)) * 2);
        S.roff[r] = (BC == PERKS_BC_ZERO && !(yin && xin)) ? kOob : off;
    }
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        const int y = y0 + r;
        S.soff[r] = (xin && y < ny) ? (uint32_t)((y * nx + x) * 2) : kOob;
        S.cidx[r] = (uint32_t)y * (uint32_t)nx + (uint32_t)x;
    }
    {
        const int hr = min(lane & 31, NY - 1);
        const int hy = y0 + hr;
        const int hyc = hy >= ny ? ny - 1 : hy;
        const int hx = (lane < 32) ? x0 - 1 : x0 + 128;
        const int hxc = hx < 0 ? 0 : (hx >= nx ? nx - 1 : hx);
        const bool hin = hx >= 0 && hx < nx && hy < ny;
        const uint32_t off = (uint32_t)((hyc * nx + hxc) * 2);
        S.hoff = (BC == PERKS_BC_ZERO && !hin) ? kOob : off;
    }

    uint32_t A[NY + 2], B[NY + 2], C[NY + 2], D[NY + 2];
    uint32_t hA, hB, hC, hD;
    S.load(zb - 1, A, hA);
    S.load(zb, B, hB);
    S.load(zb + 1, C, hC);
    for (int z = zb; z < ze; z += 4) {
        S.stage(z, A, B, hB, C, D, hD);
        if (z + 1 >= ze) break;
        S.stage(z + 1, B, C, hC, D, A, hA);
        if (z + 2 >= ze) break;
        S.stage(z + 2, C, D, hD, A, B, hB);
        if (z + 3 >= ze) break;
        S.stage(z + 3, D, A, hA, B, C, hC);
    }
}

template <int FMT, int BC, bool SR>
int launch16(const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const KCoefs& cf, int zchunk,
             uint32_t key, hipStream_t stream, PlainGeom* geom)
{
    constexpr int NY = 8;
    const int tiles_x = (nx + 127) / 128;
    const int tiles_y = (ny + 4 * NY - 1) / (4 * NY);
    const int tiles_z = (nz + zchunk - 1) / zchunk;
    const long long total = (long long)tiles_x * tiles_y * tiles_z;
    if (total > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
    if (geom) {
        geom->grid = (int)total;
        geom->block = 256;
        geom->zchunk = zchunk;
    }
    const dim3 grid((unsigned)total), block(256);
    if (cf.divisor != 0.0)
        hipLaunchKernelGGL((j3d_star1_plain16_kernel<FMT, BC, NY, true, SR>), grid, block, 0, stream, in, out, nz, ny,
                           nx, zchunk, tiles_x, tiles_y, cf, key);
    else
        hipLaunchKernelGGL((j3d_star1_plain16_kernel<FMT, BC, NY, false, SR>), grid, block, 0, stream, in, out, nz,
                           ny, nx, zchunk, tiles_x, tiles_y, cf, key);
    return hipGetLastError() == hipSuccess ? PERKS_OK : PERKS_ERR_HIP;
}

template <int FMT, int BC>
int launch16_sr(bool sr, const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const KCoefs& cf, int zchunk,
                uint32_t key, hipStream_t stream, PlainGeom* geom)
{
    return sr ? launch16<FMT, BC, true>(in, out, nz, ny, nx, cf, zchunk, key, stream, geom)
              : launch16<FMT, BC, false>(in, out, nz, ny, nx, cf, zchunk, key, stream, geom);
}

}  // namespace

namespace {

template <int FMT, int BC>
void* fn16(bool has_div, bool sr)
{
    if (has_div)
        return sr ? reinterpret_cast<void*>(&j3d_star1_plain16_kernel<FMT, BC, 8, true, true>)
                  : reinterpret_cast<void*>(&j3d_star1_plain16_kernel<FMT, BC, 8, true, false>);
    return sr ? reinterpret_cast<void*>(&j3d_star1_plain16_kernel<FMT, BC, 8, false, true>)
              : reinterpret_cast<void*>(&j3d_star1_plain16_kernel<FMT, BC, 8, false, false>);
}

template <int FMT>
void* fn16_bc(int bc, bool has_div, bool sr)
{
    return bc == PERKS_BC_ZERO ? fn16<FMT, PERKS_BC_ZERO>(has_div, sr) : fn16<FMT, PERKS_BC_CLAMP>(has_div, sr);
}

}  // namespace

void* j3d_star1_plain16_fn(int fmt, int bc, bool has_div, bool sr)
{
    if (fmt == PERKS_BF16) return fn16_bc<PERKS_BF16>(bc, has_div, sr);
    if (fmt == PERKS_F16) return fn16_bc<PERKS_F16>(bc, has_div, sr);
    return nullptr;
}

int launch_j3d_star1_plain16(int fmt, const uint16_t* in, uint16_t* out, int nz, int ny, int nx, const KCoefs& cf,
                             int bc, int zchunk, bool sr, uint32_t key, hipStream_t stream, PlainGeom* geom)
{
    if ((nx & 1) != 0) return PERKS_ERR_UNSUPPORTED_SIZE;
    if (fmt == PERKS_BF16)
        return bc == PERKS_BC_ZERO
                   ? launch16_sr<PERKS_BF16, PERKS_BC_ZERO>(sr, in, out, nz, ny, nx, cf, zchunk, key, stream, geom)
                   : launch16_sr<PERKS_BF16, PERKS_BC_CLAMP>(sr, in, out, nz, ny, nx, cf, zchunk, key, stream, geom);
    return bc == PERKS_BC_ZERO
               ? launch16_sr<PERKS_F16, PERKS_BC_ZERO>(sr, in, out, nz, ny, nx, cf, zchunk, key, stream, geom)
               : launch16_sr<PERKS_F16, PERKS_BC_CLAMP>(sr, in, out, nz, ny, nx, cf, zchunk, key, stream, geom);
}

}  // namespace perks
