// j2d_stream_bf16.hip -- instantiates the row-streaming 2D Jacobi kernels for bf16 storage.
#include "j2d_stream_kernel.h"

namespace perks {
void* j2d_stream_kernel_bf16(int shape, int bc, bool has_div, bool pair)
{
    return j2ds::sptr<j2dk::bf16s>(shape, bc, has_div, pair);
}
}  // namespace perks
