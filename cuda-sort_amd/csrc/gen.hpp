// Device input generators (gen.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cs {
hipError_t device_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
hipError_t generate(int dtype, int dist, int64_t n, int64_t count, uint64_t seed, void* out, hipStream_t s);
}
