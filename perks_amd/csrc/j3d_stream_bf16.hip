// j3d_stream_bf16.hip -- instantiates the register-streaming 3D kernels for bf16 storage.
#include "j3d_stream_kernel.h"

namespace perks {
void* j3s_kernel_bf16(int shape, int bc, int rows) { return j3s::sptr<j3g::bf16s>(shape, bc, rows); }
}  // namespace perks
