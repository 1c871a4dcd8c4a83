// Segmented (ragged) kernels for CS_I64 (see segmented.hpp).
#include "segmented.hpp"

namespace cs {
SegKernels seg_kernels_i64() { return seg_kernels<CS_I64>(); }
}  // namespace cs
