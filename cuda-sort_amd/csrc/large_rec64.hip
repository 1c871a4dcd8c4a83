// Large-array kernels (merge passes, partition) for CS_REC_I64_I32 (see large.hpp, partition.hpp).
#include "partition.hpp"

namespace cs {
LargeKernels large_kernels_rec64() { return large_kernels<CS_REC_I64_I32>(); }
PartKernels part_kernels_rec64() { return part_kernels<CS_REC_I64_I32>(); }
}  // namespace cs
