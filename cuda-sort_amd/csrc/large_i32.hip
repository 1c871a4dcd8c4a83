// Large-array kernels (merge passes, partition) for CS_I32 (see large.hpp, partition.hpp).
#include "partition.hpp"

namespace cs {
LargeKernels large_kernels_i32() { return large_kernels<CS_I32>(); }
PartKernels part_kernels_i32() { return part_kernels<CS_I32>(); }
}  // namespace cs
