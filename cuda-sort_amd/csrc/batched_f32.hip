// Batched-sort kernels for CS_F32 (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_f32(int logw, int loge) { return tile_lookup<CS_F32>(logw, loge); }
KernelInfo lds_kernel_f32() { return lds_lookup<CS_F32>(); }
}  // namespace cs
