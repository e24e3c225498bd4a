// j3d_general_bf16.hip -- instantiates the shape-generic 3D kernels for bf16 fields.
#include "j3d_general_kernel.h"

namespace perks {
void* j3g_kernel_bf16(int shape, int bc) { return j3g::kptr<j3g::bf16s>(shape, bc); }
}  // namespace perks
