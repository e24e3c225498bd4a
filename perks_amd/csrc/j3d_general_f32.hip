// j3d_general_f32.hip -- instantiates the shape-generic 3D kernels for f32 fields.
#include "j3d_general_kernel.h"

namespace perks {
void* j3g_kernel_f32(int shape, int bc) { return j3g::kptr<float>(shape, bc); }
}  // namespace perks
