// Batched-sort tile kernels for CS_F64 with W < 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_f64_lo(int logw, int loge) { return tile_lookup<CS_F64, 0, kSplitLogW>(logw, loge); }
KernelInfo lds_kernel_f64() { return lds_lookup<CS_F64>(); }
}  // namespace cs
