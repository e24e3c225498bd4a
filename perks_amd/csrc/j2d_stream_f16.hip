// j2d_stream_f16.hip -- instantiates the row-streaming 2D Jacobi kernels for f16 storage.
#include "j2d_stream_kernel.h"

namespace perks {
void* j2d_stream_kernel_f16(int shape, int bc, bool has_div, bool pair)
{
    return j2ds::sptr<j2dk::f16s>(shape, bc, has_div, pair);
}
}  // namespace perks
