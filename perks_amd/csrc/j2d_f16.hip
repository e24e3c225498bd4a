// j2d_f16.hip -- instantiates the 2D Jacobi kernels for FP16 storage.
#include "j2d_kernel.h"

namespace perks {
void* j2d_kernel_f16(int shape, int bc, bool has_div, int tile)
{
    return j2dk::kptr<j2dk::f16s>(shape, bc, has_div, tile);
}
}  // namespace perks
