// cs_sort_merge_static: the merge sort of cs_sort_merge with its scratch
// (a ping-pong buffer of n elements and one co-rank per merge tile) taken
// from a fixed device array instead of an allocation, for callers that
// cannot allocate (code under stream capture, tight serving loops).
//
// The array lives for the whole process on every device the library runs
// on.  Host code never names it directly: its address on each device is
// looked up once, with that device current, and kept per device.  Every call
// on one device shares that device's array; no word of it needs to start at
// zero (each call writes every word it reads), and calls that overlap on one
// device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/cusort_hip.h"
#include "scratch.hpp"

constexpr size_t kStaticScratchBytes = size_t(64) << 20;   // 64 MiB per device

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call there.
int static_scratch(void** p) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return int(e);
    if (dev < 0 || dev >= kMaxDevices) return CS_E_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* q = nullptr;
        e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q) {
            (void)hipGetLastError();
            return e != hipSuccess ? int(e) : CS_E_NO_DEVICE;
        }
        g_ptr[dev] = q;
    }
    *p = g_ptr[dev];
    return CS_OK;
}

bool valid_dtype(int d) { return d >= CS_I32 && d <= CS_F64; }

}  // namespace

extern "C" {

size_t cs_sort_merge_static_max_bytes(void) { return kStaticScratchBytes; }

int cs_sort_merge_static(int dtype, void* data, int64_t n, hipStream_t stream) {
    if (!valid_dtype(dtype) || n < 0) return CS_E_INVALID_VALUE;
    if (n <= 1) return CS_OK;
    if (!data) return CS_E_INVALID_VALUE;
    const size_t values = cs::scratch_round(size_t(n) * cs_dtype_size(dtype));
    const size_t need = values + cs::merge_scratch_bytes(dtype, n);
    if (need > kStaticScratchBytes) return CS_E_CAPACITY;
    void* base = nullptr;
    const int rc = static_scratch(&base);
    if (rc != CS_OK) return rc;
    char* b = static_cast<char*>(base);
    cs::Scratch tmp(stream, b + values, kStaticScratchBytes - values);
    return cs::merge_sort(dtype, data, n, data, b, tmp);
}

}  // extern "C"
