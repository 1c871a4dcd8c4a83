// Batched-sort kernels for CS_REC_I32_I32 (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_rec32(int logw, int loge) { return tile_lookup<CS_REC_I32_I32>(logw, loge); }
KernelInfo lds_kernel_rec32() { return lds_lookup<CS_REC_I32_I32>(); }
}  // namespace cs
