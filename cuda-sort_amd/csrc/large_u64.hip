// Large-array kernels (merge passes, partition) for CS_U64 (see large.hpp, partition.hpp).
#include "partition.hpp"

namespace cs {
LargeKernels large_kernels_u64() { return large_kernels<CS_U64>(); }
PartKernels part_kernels_u64() { return part_kernels<CS_U64>(); }
}  // namespace cs
