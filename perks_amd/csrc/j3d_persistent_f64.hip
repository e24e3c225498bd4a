// j3d_persistent_f64.hip -- instantiates the persistent 3D kernels for FP64 fields.
#include "j3d_persistent_kernel.h"

namespace perks {
void* j3d_persistent_kernel_f64(int bc, bool has_div, int variant, bool sr, int up)
{
    return j3dp::pick_kernel<PERKS_F64>(bc, has_div, variant, sr, up);
}
}  // namespace perks
