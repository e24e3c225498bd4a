// j3d_persistent_f32.hip -- instantiates the persistent 3D kernels for FP32 fields.
#include "j3d_persistent_kernel.h"

namespace perks {
void* j3d_persistent_kernel_f32(int bc, bool has_div, int variant, bool sr, int up)
{
    return j3dp::pick_kernel<PERKS_F32>(bc, has_div, variant, sr, up);
}
}  // namespace perks
