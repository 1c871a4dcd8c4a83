// Batched-sort tile kernels for CS_REC_I64_I32 with W < 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_rec64_lo(int logw, int loge) { return tile_lookup<CS_REC_I64_I32, 0, kSplitLogW>(logw, loge); }
KernelInfo lds_kernel_rec64() { return lds_lookup<CS_REC_I64_I32>(); }
}  // namespace cs
