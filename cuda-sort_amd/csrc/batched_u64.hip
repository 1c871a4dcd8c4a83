// Batched-sort kernels for CS_U64 (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_u64(int logw, int loge) { return tile_lookup<CS_U64>(logw, loge); }
KernelInfo lds_kernel_u64() { return lds_lookup<CS_U64>(); }
}  // namespace cs
