// j3d_stream_f64.hip -- instantiates the register-streaming 3D kernels for f64 storage.
#include "j3d_stream_kernel.h"

namespace perks {
void* j3s_kernel_f64(int shape, int bc, int rows) { return j3s::sptr<double>(shape, bc, rows); }
}  // namespace perks
