// planner.h -- host planning functions (no device calls when caps are given).
#pragma once

#include "perks_amd.h"

namespace perks {

int plan_plain(int nz, int ny, int nx, int cu_count, perks_plan_info* out);
int plan_persistent(int dtype, int nz, int ny, int nx, const perks_stencil_opts* o, const perks_device_caps* caps,
                    int occupancy_hint, perks_plan_info* out);
int min_height_persistent(int dtype, int ny, int nx, const perks_stencil_opts* o);
int min_width_y_persistent(int dtype, int nx, int ny, int bdimx, int double_tile);
// Rows of nx cells that fit in every CU's VGPR file (use_regs) plus LDS
// (use_lds): the 2D counterpart of the capacity height; 0 with neither.
int min_rows_2d_caps(int dtype, int nx, bool use_regs, bool use_lds, const perks_device_caps* caps);
int device_caps_cached(perks_device_caps* caps);

}  // namespace perks
