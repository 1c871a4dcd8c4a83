// Batched-sort kernels for CS_I32 (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_i32(int logw, int loge) { return tile_lookup<CS_I32>(logw, loge); }
KernelInfo lds_kernel_i32() { return lds_lookup<CS_I32>(); }
}  // namespace cs
