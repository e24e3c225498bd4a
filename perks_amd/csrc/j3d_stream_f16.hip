// j3d_stream_f16.hip -- instantiates the register-streaming 3D kernels for f16 storage.
#include "j3d_stream_kernel.h"

namespace perks {
void* j3s_kernel_f16(int shape, int bc, int rows) { return j3s::sptr<j3g::f16s>(shape, bc, rows); }
}  // namespace perks
