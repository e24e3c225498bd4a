// j3d_persistent_f64.hip -- instantiates the persistent 3D kernels for double.
#include "j3d_persistent_kernel.h"

namespace perks {
void* j3d_persistent_kernel_f64(int bc, bool has_div, int variant)
{
    return j3dp::pick_kernel<double>(bc, has_div, variant);
}
}  // namespace perks
