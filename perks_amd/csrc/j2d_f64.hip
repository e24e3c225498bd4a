// j2d_f64.hip -- instantiates the 2D Jacobi kernels for double.
#include "j2d_kernel.h"

namespace perks {
void* j2d_kernel_f64(int shape, int bc, bool has_div, int tile)
{
    return j2dk::kptr<double>(shape, bc, has_div, tile);
}
}  // namespace perks
