// j3d_persistent_bf16.hip -- instantiates the persistent 3D kernels for BF16
// storage with FP32 arithmetic.
#include "j3d_persistent_kernel.h"

namespace perks {
void* j3d_persistent_kernel_bf16(int bc, bool has_div, int variant, bool sr, int up)
{
    return j3dp::pick_kernel<PERKS_BF16>(bc, has_div, variant, sr, up);
}
}  // namespace perks
