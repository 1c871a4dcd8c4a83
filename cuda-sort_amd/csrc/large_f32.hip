// Large-array kernels (merge passes, partition) for CS_F32 (see large.hpp, partition.hpp).
#include "partition.hpp"

namespace cs {
LargeKernels large_kernels_f32() { return large_kernels<CS_F32>(); }
PartKernels part_kernels_f32() { return part_kernels<CS_F32>(); }
}  // namespace cs
