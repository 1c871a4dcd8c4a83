// Segmented (ragged) kernels for CS_REC_I32_I32 (see segmented.hpp).
#include "segmented.hpp"

namespace cs {
SegKernels seg_kernels_rec32() { return seg_kernels<CS_REC_I32_I32>(); }
}  // namespace cs
