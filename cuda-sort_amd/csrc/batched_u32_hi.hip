// Batched-sort tile kernels for CS_U32 with W >= 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_u32_hi(int logw, int loge) { return tile_lookup<CS_U32, kSplitLogW, 32>(logw, loge); }
}  // namespace cs
