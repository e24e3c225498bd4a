// j3d_resident.hip -- persistent 3D 7-point Jacobi with the whole domain resident in VGPRs.
//
// The reference's cached regime: kernel3d_general keeps every plane of a
// block's tile in registers (r_space) and shared memory (sm_space) across
// grid.sync()s once the domain is small enough (j3d-general-kernels.cuh:
// 258-268, prologue fill :313-380), and exchanges only the halo through
// global memory every step (the l2_cache paths, :767-1200); getminHeight
// (j3d.cu:582-612) is the height at which that holds.  Redesigned for gfx950:
//
//   * One wave owns a block of 62 columns x BY rows x BZ planes, one column
//     per lane (lanes 1..62), every cell in a VGPR for the whole solve.
//     Lanes 0 and 63 hold the neighbours' edge columns (ghost lanes), so east
//     and west come from DPP whole-wave shifts as in the streaming kernels.
//   * Every time step: load the halo (the ghost lanes of each row, the rows
//     just below and above the block, the planes just below and above it),
//     update the block in place plane by plane, store the block's shell (the
//     face rows and planes whole, the two edge columns of the inner rows) to
//     the other field buffer, and meet at the grid barrier.  Interior cells
//     never leave the chip until the final step, which stores the block whole.
//   * Cells of a block beyond a domain face (the last block along an axis may
//     overhang) hold what the boundary rule reads there at every step (0, or
//     the face cell's value), so the update needs no boundary tests; halo
//     loads beyond a face read 0 (zero rule: out-of-range buffer accesses) or
//     the clamped face cell, which is always part of some block's stored shell
//     whenever its value is used.
//   * Arithmetic is star1_update, the per-step kernel's, so the result equals
//     the CPU gold bit for bit.
//
// The grid is sized to the blocks (at most 256 CUs x occupancy workgroups of
// NW waves); a domain that needs more waves than the chip holds is refused
// (PERKS_ERR_UNSUPPORTED_SIZE): the temporal-blocking kernel serves it.
#include "j3d_resident.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "api_common.h"
#include "grid_sync.h"
#include "planner.h"
#include "j3d_common.h"

namespace perks {
namespace j3dr {

constexpr int kOwn = 62;  // columns a wave owns: lanes 1..62
constexpr int kWaves = 8; // waves per workgroup

struct ResArgs {
    int nz, ny, nx, steps;
    int nbx, nby, nbz;  // blocks per axis
    GridSync* sync;
    unsigned long long timeout;
};

// Rows per block and planes per block: FP64 8 x 6 (48 cells per lane, about
// 200 VGPRs with the halo), FP32 8 x 12.
template <typename T>
struct Shape {
    static constexpr int BY = 8;
    static constexpr int BZ = sizeof(T) == 8 ? 6 : 12;
};

// Buffer access with a scalar (wave-uniform) byte offset besides the
// per-lane one: the row's offset goes in the scalar operand, so a row costs no
// VALU address arithmetic and one field-wide descriptor serves every row.
__device__ __forceinline__ double ld(rsrc_t r, uint32_t vo, uint32_t so, double)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ float ld(rsrc_t r, uint32_t vo, uint32_t so, float)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ void st(rsrc_t r, uint32_t vo, uint32_t so, double v)
{
    using v2i = int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ void st(rsrc_t r, uint32_t vo, uint32_t so, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)vo, (int)so, 0);
}

// Addressing: a field-wide buffer descriptor per field; the plane's byte
// offset goes in the scalar offset operand (BZ + 2 wave-uniform values), the
// row and column in the per-lane offset, precomputed once per row of the
// block (BY + 2 values for loads, BY each for the ghost lanes and the two
// kinds of store).  Offsets out of range (beyond a face under the zero rule,
// lanes with nothing to store) read 0 and drop the store.
template <typename T, int BC, bool HAS_DIV>
struct Res {
    static constexpr int BY = Shape<T>::BY;
    static constexpr int BZ = Shape<T>::BZ;
    int nz, ny, nx;
    uint32_t plane_bytes, field_bytes;
    int lane, y0, z0;
    bool xdead;       // an owned lane beyond the x face
    bool has_xdead;   // the block has such lanes
    int lane_e;       // lane of column nx - 1 (clamp of the dead lanes)
    uint32_t vld[BY + 2];  // loads of rows y0-1 .. y0+BY, every lane
    uint32_t vgh[BY];      // loads of the block's rows, ghost lanes only
    uint32_t vsa[BY];      // stores of the block's rows, every owned in-domain lane
    uint32_t vse[BY];      // stores of the block's rows, the two edge columns
    T k[8];
    T v[BZ][BY];

    // Byte offset of plane zz (clamped), and whether it lies in the domain.
    __device__ __forceinline__ uint32_t zoff(int zz) const
    {
        const int zc = zz < 0 ? 0 : (zz >= nz ? nz - 1 : zz);
        return (uint32_t)zc * plane_bytes;
    }
    __device__ __forceinline__ bool zin(int zz) const { return zz >= 0 && zz < nz; }
    // Plane zz at per-lane offset vo: beyond a z face the zero rule reads 0,
    // the clamp rule the face plane.
    __device__ __forceinline__ T load(rsrc_t r, int zz, uint32_t vo) const
    {
        if (BC == PERKS_BC_ZERO && !zin(zz)) return (T)0;
        return ld(r, vo, zoff(zz), T{});
    }

    __device__ __forceinline__ void load_block(const T* f)
    {
        const rsrc_t r = make_rsrc(f, field_bytes);
#pragma unroll
        for (int z = 0; z < BZ; ++z)
#pragma unroll
            for (int y = 0; y < BY; ++y) v[z][y] = load(r, z0 + z, vld[y + 1]);
    }

    __device__ __forceinline__ void load_ghosts(const T* f)
    {
        const rsrc_t r = make_rsrc(f, field_bytes);
        const bool ghost = lane == 0 || lane == 63;
#pragma unroll
        for (int z = 0; z < BZ; ++z)
#pragma unroll
            for (int y = 0; y < BY; ++y) {
                const T g = load(r, z0 + z, vgh[y]);
                v[z][y] = ghost ? g : v[z][y];
            }
    }

    // One time step in place: plane z from planes z-1 (kept as `prev`, its
    // old values), z and z+1, and the halo rows/planes.
    __device__ __forceinline__ void step(const T* f)
    {
        const rsrc_t r = make_rsrc(f, field_bytes);
        T yl[BZ], yh[BZ], zh[BY], prev[BY];
#pragma unroll
        for (int z = 0; z < BZ; ++z) {
            yl[z] = load(r, z0 + z, vld[0]);
            yh[z] = load(r, z0 + z, vld[BY + 1]);
        }
#pragma unroll
        for (int y = 0; y < BY; ++y) {
            prev[y] = load(r, z0 - 1, vld[y + 1]);
            zh[y] = load(r, z0 + BZ, vld[y + 1]);
        }
        const int ydead = ny - y0;  // first block row beyond the y face
        const int zdead = nz - z0;
#pragma unroll
        for (int z = 0; z < BZ; ++z) {
            T out[BY];
#pragma unroll
            for (int y = 0; y < BY; ++y) {
                const T c = v[z][y];
                const T w = dpp_shift0<kDppWaveShr1>(c);
                const T e = dpp_shift0<kDppWaveShl1>(c);
                const T n = y + 1 < BY ? v[z][y + 1] : yh[z];
                const T s = y > 0 ? v[z][y - 1] : yl[z];
                const T t = z + 1 < BZ ? v[z + 1][y] : zh[y];
                out[y] = star1_update<T, HAS_DIV>(k, e, w, n, s, t, prev[y], c);
            }
            // cells beyond the faces take the boundary value
            if (has_xdead) {
#pragma unroll
                for (int y = 0; y < BY; ++y) {
                    if constexpr (BC == PERKS_BC_ZERO) out[y] = xdead ? (T)0 : out[y];
                    else out[y] = xdead ? read_lane(out[y], lane_e) : out[y];
                }
            }
            if (ydead < BY) {
#pragma unroll
                for (int y = 1; y < BY; ++y)
                    if (y >= ydead) out[y] = BC == PERKS_BC_ZERO ? (T)0 : out[y - 1];
            }
            if (z > 0 && z >= zdead) {
#pragma unroll
                for (int y = 0; y < BY; ++y) out[y] = BC == PERKS_BC_ZERO ? (T)0 : v[z - 1][y];
            }
#pragma unroll
            for (int y = 0; y < BY; ++y) {
                prev[y] = v[z][y];
                v[z][y] = out[y];
            }
        }
    }

    // The block's shell (whole: every cell) into field f.
    __device__ __forceinline__ void store(T* f, bool whole) const
    {
        const rsrc_t r = make_rsrc(f, field_bytes);
#pragma unroll
        for (int z = 0; z < BZ; ++z) {
            if (z0 + z >= nz) break;
#pragma unroll
            for (int y = 0; y < BY; ++y) {
                const bool face = y == 0 || y == BY - 1 || z == 0 || z == BZ - 1;
                st(r, (whole || face) ? vsa[y] : vse[y], zoff(z0 + z), v[z][y]);
            }
        }
    }
};

template <typename T, int BC, bool HAS_DIV>
__global__ void __launch_bounds__(kWaves * 64)
j3d_star1_resident_kernel(T* buf_a, T* buf_b, ResArgs a, KCoefs cf)
{
    using R = Res<T, BC, HAS_DIV>;
    R m;
    m.nz = a.nz;
    m.ny = a.ny;
    m.nx = a.nx;
    const uint32_t row_bytes = (uint32_t)a.nx * (uint32_t)sizeof(T);
    m.plane_bytes = (uint32_t)a.ny * row_bytes;
    m.field_bytes = (uint32_t)a.nz * m.plane_bytes;
    load_coefs7<T>(cf, m.k);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gw = blockIdx.x * kWaves + wave;
    const bool active = gw < a.nbx * a.nby * a.nbz;  // wave-uniform
    const int bx = gw % a.nbx, by = (gw / a.nbx) % a.nby, bz = gw / (a.nbx * a.nby);
    m.lane = lane;
    m.y0 = by * R::BY;
    m.z0 = bz * R::BZ;
    const int x = bx * kOwn - 1 + lane, nx = a.nx, ny = a.ny;
    const int xc = x < 0 ? 0 : (x >= nx ? nx - 1 : x);
    const bool xin = x >= 0 && x < nx;
    const bool ghost = lane == 0 || lane == 63;
    const bool own = !ghost && x < nx;
    m.xdead = !ghost && x >= nx;
    m.has_xdead = bx * kOwn + kOwn > nx;
    m.lane_e = min(max(nx - 1 - (bx * kOwn - 1), 0), 63);
#pragma unroll
    for (int j = 0; j < R::BY + 2; ++j) {
        const int yy = m.y0 - 1 + j;
        const int yc = yy < 0 ? 0 : (yy >= ny ? ny - 1 : yy);
        const bool in = (BC == PERKS_BC_CLAMP || xin) && (BC == PERKS_BC_CLAMP || (yy >= 0 && yy < ny));
        m.vld[j] = in ? (uint32_t)yc * row_bytes + (uint32_t)xc * (uint32_t)sizeof(T) : kOob;
    }
#pragma unroll
    for (int y = 0; y < R::BY; ++y) {
        const bool yin = m.y0 + y < ny;
        const uint32_t o = (uint32_t)(m.y0 + y) * row_bytes + (uint32_t)x * (uint32_t)sizeof(T);
        m.vgh[y] = ghost ? m.vld[y + 1] : kOob;
        m.vsa[y] = own && yin ? o : kOob;
        m.vse[y] = own && yin && (lane == 1 || lane == kOwn) ? o : kOob;
    }

    if (active) m.load_block(buf_a);
    for (int t = 0; t < a.steps; ++t) {
        const T* src = (t & 1) ? buf_b : buf_a;
        T* dst = (t & 1) ? buf_a : buf_b;
        const bool last = t + 1 == a.steps;
        if (active) {
            if (t > 0) m.load_ghosts(src);
            m.step(src);
            m.store(dst, last);
        }
        if (!last && !grid_sync(a.sync, gridDim.x, a.timeout)) return;
    }
}

template <typename T, int BC>
void* pick(bool has_div)
{
    return has_div ? reinterpret_cast<void*>(&j3d_star1_resident_kernel<T, BC, true>)
                   : reinterpret_cast<void*>(&j3d_star1_resident_kernel<T, BC, false>);
}

template <typename T>
void* pick_bc(int bc, bool has_div)
{
    return bc == PERKS_BC_ZERO ? pick<T, PERKS_BC_ZERO>(has_div) : pick<T, PERKS_BC_CLAMP>(has_div);
}

void blocks_of(int dtype, int nz, int ny, int nx, int& nbx, int& nby, int& nbz)
{
    const int BZ = dtype == PERKS_F64 ? Shape<double>::BZ : Shape<float>::BZ;
    nbx = (nx + kOwn - 1) / kOwn;
    nby = (ny + Shape<double>::BY - 1) / Shape<double>::BY;
    nbz = (nz + BZ - 1) / BZ;
}

}  // namespace j3dr

void* j3d_resident_fn(int dtype, int bc, bool has_div, int* block)
{
    if (block) *block = j3dr::kWaves * 64;
    if (dtype == PERKS_F64) return j3dr::pick_bc<double>(bc, has_div);
    if (dtype == PERKS_F32) return j3dr::pick_bc<float>(bc, has_div);
    return nullptr;
}

int plan_resident(int dtype, int nz, int ny, int nx, const perks_device_caps* caps, int occupancy,
                  perks_plan_info* out)
{
    if (dtype != PERKS_F64 && dtype != PERKS_F32) return PERKS_ERR_UNSUPPORTED;
    if (occupancy <= 0) return PERKS_ERR_NOT_LAUNCHABLE;
    const long long esz = dtype == PERKS_F64 ? 8 : 4;
    if ((long long)nz * ny * nx * esz >= (long long)kOob) return PERKS_ERR_UNSUPPORTED_SIZE;
    int nbx, nby, nbz;
    j3dr::blocks_of(dtype, nz, ny, nx, nbx, nby, nbz);
    const long long waves = (long long)caps->cu_count * occupancy * j3dr::kWaves;
    const long long blocks = (long long)nbx * nby * nbz;
    if (blocks > waves) return PERKS_ERR_UNSUPPORTED_SIZE;
    std::memset(out, 0, sizeof(*out));
    out->grid[0] = (int)((blocks + j3dr::kWaves - 1) / j3dr::kWaves);
    out->grid[1] = out->grid[2] = 1;
    out->block = j3dr::kWaves * 64;
    out->occupancy = occupancy;
    out->tile_x = j3dr::kOwn;
    out->tile_y = j3dr::Shape<double>::BY;
    out->reg_planes = dtype == PERKS_F64 ? j3dr::Shape<double>::BZ : j3dr::Shape<float>::BZ;
    out->reserved[0] = 1;
    return PERKS_OK;
}

int resident_height(int dtype, int ny, int nx, const perks_device_caps* caps, int occupancy)
{
    if (dtype != PERKS_F64 && dtype != PERKS_F32) return PERKS_ERR_UNSUPPORTED;
    if (ny <= 0 || nx <= 0 || occupancy <= 0) return PERKS_ERR_INVALID_ARG;
    int nbx, nby, nbz;
    j3dr::blocks_of(dtype, 1, ny, nx, nbx, nby, nbz);
    const long long waves = (long long)caps->cu_count * occupancy * j3dr::kWaves;
    const long long layers = waves / ((long long)nbx * nby);
    const int BZ = dtype == PERKS_F64 ? j3dr::Shape<double>::BZ : j3dr::Shape<float>::BZ;
    if (layers < 1) return PERKS_ERR_HEIGHT_TOO_SMALL;
    return (int)std::min<long long>(layers * BZ, 0x7fffffffLL);
}

int resident_occupancy(int dtype, int bc, bool has_div, int* occ)
{
    int block = 0;
    void* fn = j3d_resident_fn(dtype, bc, has_div, &block);
    if (!fn) return PERKS_ERR_UNSUPPORTED;
    PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, reinterpret_cast<const void*>(fn), block, 0));
    return PERKS_OK;
}

int run_resident(int dtype, void* d_a, void* d_b, int nz, int ny, int nx, int steps, const KCoefs& cf,
                 const perks_stencil_opts* o, hipStream_t stream, perks_run_report* rep)
{
    if (dtype != PERKS_F64 && dtype != PERKS_F32) return PERKS_ERR_UNSUPPORTED;
    perks_device_caps caps;
    int rc = device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    if (!caps.cooperative) return PERKS_ERR_NOT_LAUNCHABLE;
    const bool has_div = cf.divisor != 0.0;
    int block = 0;
    void* fn = j3d_resident_fn(dtype, o->boundary, has_div, &block);
    if (!fn) return PERKS_ERR_UNSUPPORTED;
    int occ = 0;
    PERKS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(fn), block, 0));
    occ = std::min(occ, std::max(1, kCoopMaxWavesPerCu / j3dr::kWaves));
    if (o->blkpsm > 0 && o->blkpsm < occ) occ = o->blkpsm;
    perks_plan_info pl;
    rc = plan_resident(dtype, nz, ny, nx, &caps, occ, &pl);
    if (rc != PERKS_OK) return rc;

    SyncLease lease;
    rc = lease.acquire();
    if (rc != PERKS_OK) return rc;

    j3dr::ResArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nz = nz;
    a.ny = ny;
    a.nx = nx;
    a.steps = steps;
    j3dr::blocks_of(dtype, nz, ny, nx, a.nbx, a.nby, a.nbz);
    a.sync = lease.gs;
    a.timeout = grid_sync_timeout();
    KCoefs kc = cf;
    void* args[] = {&d_a, &d_b, &a, &kc};

    hipEvent_t ev0, ev1;
    PERKS_HIP_TRY(hipEventCreate(&ev0));
    PERKS_HIP_TRY(hipEventCreate(&ev1));
    PERKS_HIP_TRY(hipEventRecord(ev0, stream));
    if (steps > 0) PERKS_HIP_TRY(launch_persistent(fn, dim3(pl.grid[0]), dim3(pl.block), args, 0, stream));
    PERKS_HIP_TRY(hipEventRecord(ev1, stream));
    PERKS_HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0.f;
    PERKS_HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    rc = grid_sync_check(lease.gs, stream);
    if (rc != PERKS_OK) return rc;
    if (rep) {
        rep->elapsed_ms = ms;
        rep->gcells_per_s = ms > 0 ? (double)steps * nz * ny * (double)nx / ms / 1e6 : 0.0;
        rep->grid[0] = pl.grid[0];
        rep->grid[1] = rep->grid[2] = 1;
        rep->block = pl.block;
        rep->occupancy = pl.occupancy;
        rep->result_in_b = steps % 2;
        rep->reserved[0] = 1;
        rep->reserved[1] = steps;
    }
    return PERKS_OK;
}

}  // namespace perks
