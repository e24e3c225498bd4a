// j3d_persistent_f32.hip -- instantiates the persistent 3D kernels for float.
#include "j3d_persistent_kernel.h"

namespace perks {
void* j3d_persistent_kernel_f32(int bc, bool has_div, int variant)
{
    return j3dp::pick_kernel<float>(bc, has_div, variant);
}
}  // namespace perks
