// Batched-sort kernels for CS_I64 (see batched.hpp).
#include "batched.hpp"

namespace cs {
KernelInfo tile_kernel_i64(int logw, int loge) { return tile_lookup<CS_I64>(logw, loge); }
KernelInfo lds_kernel_i64() { return lds_lookup<CS_I64>(); }
}  // namespace cs
