// j2d_stream_f32.hip -- instantiates the row-streaming 2D Jacobi kernels for f32 storage.
#include "j2d_stream_kernel.h"

namespace perks {
void* j2d_stream_kernel_f32(int shape, int bc, bool has_div, bool pair)
{
    return j2ds::sptr<float>(shape, bc, has_div, pair);
}
}  // namespace perks
