// Batched-sort tile kernels for CS_F64 with W >= 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_f64_hi(int logw, int loge) { return tile_lookup<CS_F64, kSplitLogW, 32>(logw, loge); }
}  // namespace cs
