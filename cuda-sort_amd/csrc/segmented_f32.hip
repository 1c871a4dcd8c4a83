// Segmented (ragged) kernels for CS_F32 (see segmented.hpp).
#include "segmented.hpp"

namespace cs {
SegKernels seg_kernels_f32() { return seg_kernels<CS_F32>(); }
}  // namespace cs
