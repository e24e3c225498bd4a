// j3d_plain.hip -- plain per-step 3D Jacobi kernel (one launch per time step).
//
// Replaces the reference's kernel3d_baseline (stencil/3dstencil/j3d-baseline.cu:21-191),
// launched once per step with a host-side ping-pong (j3d.cu:728-740).
//
// Mapping (gfx950, wave64): a workgroup is 4 waves; each wave owns a strip of
// 64 x-columns (one per lane) by NY y-rows and marches along z through the
// workgroup's z-chunk.  Every lane keeps, in VGPRs, the NY+2 rows (the strip
// plus one halo row on each side) of planes z and z+1, the NY rows of plane
// z-1, and prefetches plane z+2 while it computes plane z.  East/west
// neighbours come from the adjacent lane through DPP whole-wave shifts; the two
// tile-edge columns of each row arrive with one extra load whose lanes r and
// 32+r fetch row r's west and east halo, then read_lane broadcasts them.
// Global traffic per cell: one read of the cell (plus 2/NY halo rows, mostly
// L2 hits) and one write.
#include "j3d_kernels.h"

namespace perks {

namespace {

template <typename T, int BC, int NY>
struct PlaneLoader {
    const T* __restrict__ in;
    int nz;
    size_t plane;
    int roff[NY + 2];  // row offsets of rows y0-1 .. y0+NY (clamped)
    bool rzero[NY + 2];
    int xl;            // clamped column of this lane
    bool xzero;
    int hoff;          // halo element of this lane (row hr, column west/east)
    bool hzero;

    __device__ __forceinline__ void load(int zz, T (&v)[NY + 2], T& h) const
    {
        if (zz < 0 || zz >= nz) {
            if (BC == PERKS_BC_ZERO) {
#pragma unroll
                for (int r = 0; r < NY + 2; ++r) v[r] = (T)0;
                h = (T)0;
                return;
            }
            zz = zz < 0 ? 0 : nz - 1;
        }
        const T* __restrict__ base = in + (size_t)zz * plane;
#pragma unroll
        for (int r = 0; r < NY + 2; ++r) {
            T x = base[roff[r] + xl];
            if (BC == PERKS_BC_ZERO) x = (rzero[r] || xzero) ? (T)0 : x;
            v[r] = x;
        }
        T hv = base[hoff];
        if (BC == PERKS_BC_ZERO) hv = hzero ? (T)0 : hv;
        h = hv;
    }
};

template <typename T, int BC, int NY, bool HAS_DIV>
__global__ void __launch_bounds__(256)
j3d_star1_plain_kernel(const T* __restrict__ in, T* __restrict__ out, int nz, int ny, int nx, int zchunk,
                       int tiles_x, int tiles_y, KCoefs cf)
{
    T k[8];
    load_coefs7<T>(cf, k);

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int total = gridDim.x;
    const int tile = xcd_tile_id(blockIdx.x, total);
    const int bx = tile % tiles_x;
    const int by = (tile / tiles_x) % tiles_y;
    const int bz = tile / (tiles_x * tiles_y);

    const int x0 = bx * 64;
    const int y0 = (by * 4 + wave) * NY;
    const int zb = bz * zchunk;
    const int ze = min(zb + zchunk, nz);
    const int x = x0 + lane;
    const bool xin = x < nx;

    PlaneLoader<T, BC, NY> L;
    L.in = in;
    L.nz = nz;
    L.plane = (size_t)ny * nx;
    L.xl = xin ? x : nx - 1;
    L.xzero = !xin;
#pragma unroll
    for (int r = 0; r < NY + 2; ++r) {
        const int yy = y0 + r - 1;
        const int yc = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
        L.roff[r] = yc * nx;
        L.rzero[r] = (yy < 0 || yy >= ny);
    }
    {
        const int hr = min(lane & 31, NY - 1);
        const int hy = y0 + hr;
        const int hyc = hy >= ny ? ny - 1 : hy;
        const int hx = (lane < 32) ? x0 - 1 : x0 + 64;
        const int hxc = hx < 0 ? 0 : (hx >= nx ? nx - 1 : hx);
        L.hoff = hyc * nx + hxc;
        L.hzero = (hx < 0 || hx >= nx || hy >= ny);
    }

    T bel[NY], cur[NY + 2], up[NY + 2], nxt[NY + 2];
    T hcur, hup, hnxt, hdummy;
    {
        T tmp[NY + 2];
        L.load(zb - 1, tmp, hdummy);
#pragma unroll
        for (int r = 0; r < NY; ++r) bel[r] = tmp[r + 1];
    }
    L.load(zb, cur, hcur);
    L.load(zb + 1, up, hup);

    for (int z = zb; z < ze; ++z) {
        if (z + 1 < ze) L.load(z + 2, nxt, hnxt);
        T* __restrict__ orow = out + (size_t)z * L.plane + x;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const T hw = read_lane(hcur, r);
            const T he = read_lane(hcur, 32 + r);
            const T c = cur[r + 1];
            const T e = dpp_move<kDppWaveShl1>(c, he);
            const T w = dpp_move<kDppWaveShr1>(c, hw);
            const T v = star1_update<T, HAS_DIV>(k, e, w, cur[r + 2], cur[r], up[r + 1], bel[r], c);
            const int y = y0 + r;
            if (xin && y < ny) orow[(size_t)y * nx] = v;
        }
#pragma unroll
        for (int r = 0; r < NY; ++r) bel[r] = cur[r + 1];
#pragma unroll
        for (int r = 0; r < NY + 2; ++r) {
            cur[r] = up[r];
            up[r] = nxt[r];
        }
        hcur = hup;
        hup = hnxt;
    }
}

template <typename T, int BC, int NY>
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
        hipLaunchKernelGGL((j3d_star1_plain_kernel<T, BC, NY, true>), grid, block, 0, stream, in, out, nz, ny, nx,
                           zchunk, tiles_x, tiles_y, cf);
    else
        hipLaunchKernelGGL((j3d_star1_plain_kernel<T, BC, NY, false>), grid, block, 0, stream, in, out, nz, ny,
                           nx, zchunk, tiles_x, tiles_y, cf);
    return hipGetLastError() == hipSuccess ? PERKS_OK : PERKS_ERR_HIP;
}

}  // namespace

// Default z-chunk: enough workgroups to give every CU several (>= 8 per CU
// when the volume allows), but chunks no shorter than 16 planes so the two
// re-read planes at each chunk start stay a small overhead.
int plain_default_zchunk(int nz, int ny, int nx, int cu_count, int ny_per_thread)
{
    const long long tiles_xy = (long long)((nx + 63) / 64) * ((ny + 4 * ny_per_thread - 1) / (4 * ny_per_thread));
    const long long want = 8LL * cu_count;
    long long chunks = (want + tiles_xy - 1) / tiles_xy;
    if (chunks < 1) chunks = 1;
    int zc = (int)((nz + chunks - 1) / chunks);
    if (zc < 16) zc = 16;
    if (zc > nz) zc = nz;
    return zc < 1 ? 1 : zc;
}

template <typename T>
int launch_j3d_star1_plain(const T* in, T* out, int nz, int ny, int nx, const KCoefs& cf, int bc, int zchunk,
                           hipStream_t stream, PlainGeom* geom)
{
    const bool has_div = cf.divisor != 0.0;
    if (bc == PERKS_BC_ZERO) return launch_plain_ny<T, PERKS_BC_ZERO, 8>(in, out, nz, ny, nx, cf, has_div, zchunk, stream, geom);
    return launch_plain_ny<T, PERKS_BC_CLAMP, 8>(in, out, nz, ny, nx, cf, has_div, zchunk, stream, geom);
}

template int launch_j3d_star1_plain<double>(const double*, double*, int, int, int, const KCoefs&, int, int,
                                            hipStream_t, PlainGeom*);
template int launch_j3d_star1_plain<float>(const float*, float*, int, int, int, const KCoefs&, int, int,
                                           hipStream_t, PlainGeom*);

}  // namespace perks
