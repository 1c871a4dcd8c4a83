// Large-array kernels (merge passes, partition) for CS_I64 (see large.hpp, partition.hpp).
#include "partition.hpp"

namespace cs {
LargeKernels large_kernels_i64() { return large_kernels<CS_I64>(); }
PartKernels part_kernels_i64() { return part_kernels<CS_I64>(); }
}  // namespace cs
