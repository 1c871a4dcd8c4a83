// Segmented (ragged) kernels for CS_U64 (see segmented.hpp).
#include "segmented.hpp"

namespace cs {
SegKernels seg_kernels_u64() { return seg_kernels<CS_U64>(); }
}  // namespace cs
