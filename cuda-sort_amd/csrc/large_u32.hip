// Large-array kernels (merge passes, partition) for CS_U32 (see large.hpp, partition.hpp).
#include "partition.hpp"

namespace cs {
LargeKernels large_kernels_u32() { return large_kernels<CS_U32>(); }
PartKernels part_kernels_u32() { return part_kernels<CS_U32>(); }
}  // namespace cs
