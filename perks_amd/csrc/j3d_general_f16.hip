// j3d_general_f16.hip -- instantiates the shape-generic 3D kernels for f16 fields.
#include "j3d_general_kernel.h"

namespace perks {
void* j3g_kernel_f16(int shape, int bc) { return j3g::kptr<j3g::f16s>(shape, bc); }
}  // namespace perks
