// j3d_stream_f32.hip -- instantiates the register-streaming 3D kernels for f32 storage.
#include "j3d_stream_kernel.h"

namespace perks {
void* j3s_kernel_f32(int shape, int bc, int rows) { return j3s::sptr<float>(shape, bc, rows); }
}  // namespace perks
