// Batched-sort tile kernels for CS_I32 with W < 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_i32_lo(int logw, int loge) { return tile_lookup<CS_I32, 0, kSplitLogW>(logw, loge); }
KernelInfo lds_kernel_i32() { return lds_lookup<CS_I32>(); }
}  // namespace cs
