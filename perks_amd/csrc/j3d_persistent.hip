// j3d_persistent.hip -- persistent 3D Jacobi (placeholder until the kernel lands).
#include "j3d_persistent.h"

namespace perks {

template <typename T>
int run_persistent(T*, T*, int, int, int, int, const KCoefs&, const perks_stencil_opts*, hipStream_t,
                   perks_run_report*)
{
    return PERKS_ERR_UNSUPPORTED;
}

template int run_persistent<double>(double*, double*, int, int, int, int, const KCoefs&, const perks_stencil_opts*,
                                    hipStream_t, perks_run_report*);
template int run_persistent<float>(float*, float*, int, int, int, int, const KCoefs&, const perks_stencil_opts*,
                                   hipStream_t, perks_run_report*);

int release_persistent_scratch() { return PERKS_OK; }

}  // namespace perks

#include "planner.h"

namespace perks {
int plan_persistent(int, int, int, int, const perks_stencil_opts*, const perks_device_caps*, int, perks_plan_info*)
{
    return PERKS_ERR_UNSUPPORTED;
}
int min_height_persistent(int, int, int, const perks_stencil_opts*) { return PERKS_ERR_UNSUPPORTED; }
int min_width_y_persistent(int, int, int, int, int) { return PERKS_ERR_UNSUPPORTED; }
}  // namespace perks
