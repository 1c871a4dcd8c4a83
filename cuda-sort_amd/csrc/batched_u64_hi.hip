// Batched-sort tile kernels for CS_U64 with W >= 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_u64_hi(int logw, int loge) { return tile_lookup<CS_U64, kSplitLogW, 32>(logw, loge); }
}  // namespace cs
