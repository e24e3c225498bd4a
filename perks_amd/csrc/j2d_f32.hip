// j2d_f32.hip -- instantiates the 2D Jacobi kernels for float.
#include "j2d_kernel.h"

namespace perks {
void* j2d_kernel_f32(int shape, int bc, bool has_div, int tile)
{
    return j2dk::kptr<float>(shape, bc, has_div, tile);
}
}  // namespace perks
