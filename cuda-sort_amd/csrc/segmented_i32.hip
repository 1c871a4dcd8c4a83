// Segmented (ragged) kernels for CS_I32 (see segmented.hpp).
#include "segmented.hpp"

namespace cs {
SegKernels seg_kernels_i32() { return seg_kernels<CS_I32>(); }
}  // namespace cs
