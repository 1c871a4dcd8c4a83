// Batched-sort tile kernels for CS_U32 with W < 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_u32_lo(int logw, int loge) { return tile_lookup<CS_U32, 0, kSplitLogW>(logw, loge); }
KernelInfo lds_kernel_u32() { return lds_lookup<CS_U32>(); }
}  // namespace cs
