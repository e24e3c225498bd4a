// j3d_plain.hip -- plain per-step 3D Jacobi kernel (one launch per time step).
//
// Replaces the reference's kernel3d_baseline (stencil/3dstencil/j3d-baseline.cu:21-191),
// launched once per step with a host-side ping-pong (j3d.cu:728-740).
//
// Mapping (gfx950, wave64): a workgroup is 4 waves; each wave owns a strip of
// 64 x-columns (one per lane) by NY y-rows and marches along z through the
// workgroup's z-chunk.  Every lane keeps, in VGPRs, the NY+2 rows (the strip
// plus one halo row on each side) of planes z-1, z and z+1, and prefetches
// plane z+2 while it computes plane z.  The four plane buffers form a ring
// that the loop body walks through four times per iteration, so the window
// moves by renaming registers instead of copying them.  East/west neighbours
// come from the adjacent lane through DPP whole-wave shifts; the two
// tile-edge columns of each row arrive with one extra load whose lanes r and
// 32+r fetch row r's west and east halo, then read_lane broadcasts them.
// Global traffic per cell: one read of the cell (plus 2/NY halo rows, mostly
// L2 hits) and one write.
#include "j3d_kernels.h"

namespace perks {

namespace {

template <typename T, int BC, int NY, bool HAS_DIV, bool NT_STORE>
struct Star1Plain {
    const T* __restrict__ in;
    T* __restrict__ out;
    int nz;
    int ze;
    size_t plane;            // elements per plane
    uint32_t plane_bytes;
    uint32_t roff[NY + 2];   // byte offsets of rows y0-1 .. y0+NY in a plane (kOob: zero rule)
    uint32_t soff[NY];       // byte offsets of this lane's output cells (kOob: not stored)
    uint32_t hoff;           // this lane's halo element (row hr, west or east column)
    T k[8];

    // Issues the loads of plane zz into v / h; never waits.  Outside [0, nz)
    // the clamp rule reads the edge plane and the zero rule a size-0
    // descriptor, whose loads return 0.
    __device__ __forceinline__ void load(int zz, T (&v)[NY + 2], T& h) const
    {
        uint32_t bytes = plane_bytes;
        int zc = zz;
        if (zz < 0 || zz >= nz) {
            zc = zz < 0 ? 0 : nz - 1;
            if (BC == PERKS_BC_ZERO) bytes = 0;
        }
        const rsrc_t rs = make_rsrc(in + (size_t)zc * plane, bytes);
#pragma unroll
        for (int r = 0; r < NY + 2; ++r) v[r] = buf_load(rs, roff[r], T{});
        h = buf_load(rs, hoff, T{});
    }

    __device__ __forceinline__ void compute(int z, const T (&bel)[NY + 2], const T (&cur)[NY + 2], T hcur,
                                            const T (&up)[NY + 2]) const
    {
        const rsrc_t rs = make_rsrc(out + (size_t)z * plane, plane_bytes);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const T hw = read_lane(hcur, r);
            const T he = read_lane(hcur, 32 + r);
            const T c = cur[r + 1];
            const T e = dpp_move<kDppWaveShl1>(c, he);
            const T w = dpp_move<kDppWaveShr1>(c, hw);
            const T v = star1_update<T, HAS_DIV>(k, e, w, cur[r + 2], cur[r], up[r + 1], bel[r + 1], c);
            buf_store<NT_STORE ? 2 : 0>(rs, soff[r], v);
        }
    }

    // One stage of the ring: prefetch plane z+2 into `nxt`, compute plane z.
    __device__ __forceinline__ void stage(int z, const T (&bel)[NY + 2], const T (&cur)[NY + 2], T hcur,
                                          const T (&up)[NY + 2], T (&nxt)[NY + 2], T& hnxt) const
    {
        load(z + 2, nxt, hnxt);
        compute(z, bel, cur, hcur, up);
    }
};

template <typename T, int BC, int NY, bool HAS_DIV, bool NT_STORE>
__global__ void __launch_bounds__(256)
j3d_star1_plain_kernel(const T* __restrict__ in, T* __restrict__ out, int nz, int ny, int nx, int zchunk,
                       int tiles_x, int tiles_y, KCoefs cf)
{
    Star1Plain<T, BC, NY, HAS_DIV, NT_STORE> S;
    load_coefs7<T>(cf, S.k);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tile = xcd_tile_id(blockIdx.x, gridDim.x);
    const int bx = tile % tiles_x;
    const int by = (tile / tiles_x) % tiles_y;
    const int bz = tile / (tiles_x * tiles_y);

    const int x0 = bx * 64;
    const int y0 = (by * 4 + wave) * NY;
    const int zb = bz * zchunk;
    S.in = in;
    S.out = out;
    S.nz = nz;
    S.ze = min(zb + zchunk, nz);
    S.plane = (size_t)ny * nx;
    S.plane_bytes = (uint32_t)(S.plane * sizeof(T));
    const int x = x0 + lane;
    const bool xin = x < nx;
    // Clamp rule: columns past the edge read the edge column (so the last
    // valid lane's east neighbour, taken from the next lane, is itself).
    const int xl = xin ? x : nx - 1;
#pragma unroll
    for (int r = 0; r < NY + 2; ++r) {
        const int yy = y0 + r - 1;
        const bool yin = yy >= 0 && yy < ny;
        const int yc = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
        const uint32_t off = (uint32_t)((yc * nx + xl) * (int)sizeof(T));
        S.roff[r] = (BC == PERKS_BC_ZERO && !(yin && xin)) ? kOob : off;
    }
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        const int y = y0 + r;
        S.soff[r] = (xin && y < ny) ? (uint32_t)((y * nx + x) * (int)sizeof(T)) : kOob;
    }
    {
        const int hr = min(lane & 31, NY - 1);
        const int hy = y0 + hr;
        const int hyc = hy >= ny ? ny - 1 : hy;
        const int hx = (lane < 32) ? x0 - 1 : x0 + 64;
        const int hxc = hx < 0 ? 0 : (hx >= nx ? nx - 1 : hx);
        const bool hin = hx >= 0 && hx < nx && hy < ny;
        const uint32_t off = (uint32_t)((hyc * nx + hxc) * (int)sizeof(T));
        S.hoff = (BC == PERKS_BC_ZERO && !hin) ? kOob : off;
    }

    T A[NY + 2], B[NY + 2], C[NY + 2], D[NY + 2];
    T hA, hB, hC, hD;
    S.load(zb - 1, A, hA);
    S.load(zb, B, hB);
    S.load(zb + 1, C, hC);
    for (int z = zb; z < S.ze; z += 4) {
        S.stage(z, A, B, hB, C, D, hD);
        if (z + 1 >= S.ze) break;
        S.stage(z + 1, B, C, hC, D, A, hA);
        if (z + 2 >= S.ze) break;
        S.stage(z + 2, C, D, hD, A, B, hB);
        if (z + 3 >= S.ze) break;
        S.stage(z + 3, D, A, hA, B, C, hC);
    }
}

template <typename T, int BC, int NY, bool NT>
int launch_plain_ny(const T* in, T* out, int nz, int ny, int nx, const KCoefs& cf, bool has_div, int zchunk,
                    hipStream_t stream, PlainGeom* geom)
{
    const int tiles_x = (nx + 63) / 64;
    const int tiles_y = (ny + 4 * NY - 1) / (4 * NY);
    const int tiles_z = (nz + zchunk - 1) / zchunk;
    const long long total = (long long)tiles_x * tiles_y * tiles_z;
    if (total > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
    if (geom) {
        geom->grid = (int)total;
        geom->block = 256;
        geom->zchunk = zchunk;
    }
    dim3 grid((unsigned)total), block(256);
    if (has_div)
        hipLaunchKernelGGL((j3d_star1_plain_kernel<T, BC, NY, true, NT>), grid, block, 0, stream, in, out, nz, ny,
                           nx, zchunk, tiles_x, tiles_y, cf);
    else
        hipLaunchKernelGGL((j3d_star1_plain_kernel<T, BC, NY, false, NT>), grid, block, 0, stream, in, out, nz, ny,
                           nx, zchunk, tiles_x, tiles_y, cf);
    return hipGetLastError() == hipSuccess ? PERKS_OK : PERKS_ERR_HIP;
}

}  // namespace

template <typename T>
int launch_j3d_star1_plain(const T* in, T* out, int nz, int ny, int nx, const KCoefs& cf, int bc, int zchunk,
                           hipStream_t stream, PlainGeom* geom, int variant)
{
    const bool has_div = cf.divisor != 0.0;
    const bool nt = (variant & 1) != 0;
    if (bc == PERKS_BC_ZERO) {
        if (nt) return launch_plain_ny<T, PERKS_BC_ZERO, 8, true>(in, out, nz, ny, nx, cf, has_div, zchunk, stream, geom);
        return launch_plain_ny<T, PERKS_BC_ZERO, 8, false>(in, out, nz, ny, nx, cf, has_div, zchunk, stream, geom);
    }
    if (nt) return launch_plain_ny<T, PERKS_BC_CLAMP, 8, true>(in, out, nz, ny, nx, cf, has_div, zchunk, stream, geom);
    return launch_plain_ny<T, PERKS_BC_CLAMP, 8, false>(in, out, nz, ny, nx, cf, has_div, zchunk, stream, geom);
}

namespace {

template <typename T, int BC>
void* plain_fn(bool has_div, bool nt)
{
    if (has_div)
        return nt ? reinterpret_cast<void*>(&j3d_star1_plain_kernel<T, BC, 8, true, true>)
                  : reinterpret_cast<void*>(&j3d_star1_plain_kernel<T, BC, 8, true, false>);
    return nt ? reinterpret_cast<void*>(&j3d_star1_plain_kernel<T, BC, 8, false, true>)
              : reinterpret_cast<void*>(&j3d_star1_plain_kernel<T, BC, 8, false, false>);
}

template <typename T>
void* plain_fn_bc(int bc, bool has_div, bool nt)
{
    return bc == PERKS_BC_ZERO ? plain_fn<T, PERKS_BC_ZERO>(has_div, nt) : plain_fn<T, PERKS_BC_CLAMP>(has_div, nt);
}

}  // namespace

void* j3d_star1_plain_fn(int dtype, int bc, bool has_div, bool nontemporal)
{
    if (dtype == PERKS_F64) return plain_fn_bc<double>(bc, has_div, nontemporal);
    if (dtype == PERKS_F32) return plain_fn_bc<float>(bc, has_div, nontemporal);
    return nullptr;
}

template int launch_j3d_star1_plain<double>(const double*, double*, int, int, int, const KCoefs&, int, int,
                                            hipStream_t, PlainGeom*, int);
template int launch_j3d_star1_plain<float>(const float*, float*, int, int, int, const KCoefs&, int, int,
                                           hipStream_t, PlainGeom*, int);

}  // namespace perks
