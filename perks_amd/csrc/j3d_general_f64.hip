// j3d_general_f64.hip -- instantiates the shape-generic 3D kernels for f64 fields.
#include "j3d_general_kernel.h"

namespace perks {
void* j3g_kernel_f64(int shape, int bc) { return j3g::kptr<double>(shape, bc); }
}  // namespace perks
