// j2d_bf16.hip -- instantiates the 2D Jacobi kernels for BF16 storage.
#include "j2d_kernel.h"

namespace perks {
void* j2d_kernel_bf16(int shape, int bc, bool has_div, int tile)
{
    return j2dk::kptr<j2dk::bf16s>(shape, bc, has_div, tile);
}
}  // namespace perks
