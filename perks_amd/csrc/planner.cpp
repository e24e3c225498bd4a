// planner.cpp -- pure host planning for the 3D Jacobi path.
//
// The reference plans inside j3d_iterative (stencil/3dstencil/j3d.cu:489-612):
// occupancy -> persistent grid, LDS budget -> cached planes, refusals -1/-3/-4/-5.
// Here planning is a function of (sizes, options, device caps, occupancy) so
// that every refusal can be exercised on a machine without a GPU.
#include <cstring>

#include "j3d_general.h"
#include "perks_amd.h"
#include "perks_shapes.h"
#include "planner.h"

namespace perks {

int plan_plain(int nz, int ny, int nx, int cu_count, perks_plan_info* out)
{
    const int ny_per_thread = 8;
    const int tiles_x = (nx + 63) / 64;
    const int tiles_y = (ny + 4 * ny_per_thread - 1) / (4 * ny_per_thread);
    const long long tiles_xy = (long long)tiles_x * tiles_y;
    const long long want = 8LL * cu_count;
    long long chunks = (want + tiles_xy - 1) / tiles_xy;
    if (chunks < 1) chunks = 1;
    int zc = (int)((nz + chunks - 1) / chunks);
    if (zc < 16) zc = 16;
    if (zc > nz) zc = nz;
    if (zc < 1) zc = 1;
    const long long total = tiles_xy * ((nz + zc - 1) / zc);
    if (total > 0x7fffffffLL) return PERKS_ERR_UNSUPPORTED_SIZE;
    std::memset(out, 0, sizeof(*out));
    out->grid[0] = (int)total;
    out->grid[1] = 1;
    out->grid[2] = 1;
    out->block = 256;
    out->tile_x = 64;
    out->tile_y = 4 * ny_per_thread;
    out->reserved[0] = zc;
    return PERKS_OK;
}

int min_rows_2d_caps(int dtype, int nx, bool use_regs, bool use_lds, const perks_device_caps* caps)
{
    if (nx <= 0 || !caps) return PERKS_ERR_INVALID_ARG;
    const long long esz = dtype == PERKS_F64 ? 8 : (dtype == PERKS_F32 ? 4 : 2);
    const long long per_cu = (use_lds ? (long long)caps->lds_per_cu : 0) +
                             (use_regs ? (long long)caps->vgprs_per_simd * caps->simds_per_cu * caps->warp_size * 4LL
                                       : 0);
    const long long rows = (long long)caps->cu_count * per_cu / (nx * esz);
    return (int)(rows < 0x7fffffffLL ? rows : 0x7fffffffLL);
}

}  // namespace perks

extern "C" int perks_j3d_plan(int dtype, int height, int width_y, int width_x, const perks_stencil_opts* opts,
                              const perks_device_caps* caps, int occupancy_hint, perks_plan_info* out)
{
    if (!opts || !out || height <= 0 || width_y <= 0 || width_x <= 0) return PERKS_ERR_INVALID_ARG;
    if (dtype != PERKS_F64 && dtype != PERKS_F32 && dtype != PERKS_F16 && dtype != PERKS_BF16)
        return PERKS_ERR_UNSUPPORTED;
    perks_device_caps local;
    if (!caps) {
        int dev = 0;
        (void)dev;
        int rc = perks_query_device(0, &local);
        if (rc != PERKS_OK) return rc;
        caps = &local;
    }
    if (opts->shape != PERKS_SHAPE_3D_STAR_R1 || opts->tune[5] == 1)
        return perks::j3g_plan(dtype, height, width_y, width_x, opts, caps, occupancy_hint, out);
    if (opts->method == PERKS_METHOD_PLAIN) return perks::plan_plain(height, width_y, width_x, caps->cu_count, out);
    return perks::plan_persistent(dtype, height, width_y, width_x, opts, caps, occupancy_hint, out);
}

extern "C" int perks_j3d_min_height(int dtype, int width_y, int width_x, const perks_stencil_opts* opts)
{
    if (!opts || width_y <= 0 || width_x <= 0) return PERKS_ERR_INVALID_ARG;
    return perks::min_height_persistent(dtype, width_y, width_x, opts);
}

extern "C" int perks_get_min_width_y_2d(int dtype, int width_x, int bdimx, int registers, int use_lds, int blkpsm,
                                        int double_tile)
{
    (void)bdimx;
    (void)blkpsm;
    (void)double_tile;
    if (width_x <= 0 || registers < 0) return PERKS_ERR_INVALID_ARG;
    if (dtype != PERKS_F64 && dtype != PERKS_F32 && dtype != PERKS_F16 && dtype != PERKS_BF16)
        return PERKS_ERR_UNSUPPORTED;
    perks_device_caps caps;
    const int rc = perks::device_caps_cached(&caps);
    if (rc != PERKS_OK) return rc;
    return perks::min_rows_2d_caps(dtype, width_x, registers > 0, use_lds != 0, &caps);
}

extern "C" int perks_get_min_width_y(int dtype, int width_x, int width_y, int bdimx, int double_tile)
{
    if (width_x <= 0 || width_y <= 0) return PERKS_ERR_INVALID_ARG;
    return perks::min_width_y_persistent(dtype, width_x, width_y, bdimx, double_tile);
}
