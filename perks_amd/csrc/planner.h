// planner.h -- host planning functions (no device calls when caps are given).
#pragma once

#include "perks_amd.h"

namespace perks {

int plan_plain(int nz, int ny, int nx, int cu_count, perks_plan_info* out);
int plan_persistent(int dtype, int nz, int ny, int nx, const perks_stencil_opts* o, const perks_device_caps* caps,
                    int occupancy_hint, perks_plan_info* out);
int min_height_persistent(int dtype, int ny, int nx, const perks_stencil_opts* o);
int min_width_y_persistent(int dtype, int nx, int ny, int bdimx, int double_tile);

}  // namespace perks
