// Large-array kernels (merge passes, partition) for CS_F64 (see large.hpp, partition.hpp).
#include "partition.hpp"

namespace cs {
LargeKernels large_kernels_f64() { return large_kernels<CS_F64>(); }
PartKernels part_kernels_f64() { return part_kernels<CS_F64>(); }
}  // namespace cs
