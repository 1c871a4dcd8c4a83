// Batched-sort tile kernels for CS_F32 with W < 2^kSplitLogW (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_f32_lo(int logw, int loge) { return tile_lookup<CS_F32, 0, kSplitLogW>(logw, loge); }
KernelInfo lds_kernel_f32() { return lds_lookup<CS_F32>(); }
}  // namespace cs
