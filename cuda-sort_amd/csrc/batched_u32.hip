// Batched-sort kernels for CS_U32 (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_u32(int logw, int loge) { return tile_lookup<CS_U32>(logw, loge); }
KernelInfo lds_kernel_u32() { return lds_lookup<CS_U32>(); }
}  // namespace cs
