// j3d_persistent_f16.hip -- instantiates the persistent 3D kernels for FP16
// storage with FP32 arithmetic.
#include "j3d_persistent_kernel.h"

namespace perks {
void* j3d_persistent_kernel_f16(int bc, bool has_div, int variant, bool sr, int up)
{
    return j3dp::pick_kernel<PERKS_F16>(bc, has_div, variant, sr, up);
}
}  // namespace perks
