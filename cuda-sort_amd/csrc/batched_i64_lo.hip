// Batched-sort tile kernels for CS_I64 with W < 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_i64_lo(int logw, int loge) { return tile_lookup<CS_I64, 0, kSplitLogW>(logw, loge); }
KernelInfo lds_kernel_i64() { return lds_lookup<CS_I64>(); }
}  // namespace cs
