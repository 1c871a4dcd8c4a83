// Batched-sort tile kernels for CS_I64 with W >= 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_i64_hi(int logw, int loge) { return tile_lookup<CS_I64, kSplitLogW, 32>(logw, loge); }
}  // namespace cs
