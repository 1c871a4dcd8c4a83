// Device input generators (gen.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cs {
hipError_t device_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
hipError_t stream_hold(double microseconds, hipStream_t s);
hipError_t generate(int dtype, int dist, int64_t n, int64_t count, uint64_t seed, void* out, hipStream_t s);
// Stream-ordered allocation from a per-device pool that keeps freed memory
// cached (unlimited release threshold), so that repeated large sorts do not
// pay for the pool growing back after every synchronisation.  Freed with
// hipFreeAsync.  The pool's memory stays reserved until the process ends.
hipError_t pool_malloc_async(void** p, size_t bytes, hipStream_t s);
}
