// Segmented (ragged) kernels for CS_U32 (see segmented.hpp).
#include "segmented.hpp"

namespace cs {
SegKernels seg_kernels_u32() { return seg_kernels<CS_U32>(); }
}  // namespace cs
