// Segmented (ragged) kernels for CS_F64 (see segmented.hpp).
#include "segmented.hpp"

namespace cs {
SegKernels seg_kernels_f64() { return seg_kernels<CS_F64>(); }
}  // namespace cs
