// Device-side input generators: stateless, so every element depends only on
// (seed, global index, array length) and any launch shape produces the same
// data.  Mirrors the reference's initializeBatch kernel (bench.cu:157-192),
// which generates its distributions from the mixBits hash of the global
// index, and adds the distributions of the MI355X configurations.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cusort_hip.h"
#include "gen.hpp"

namespace cs {
namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint32_t mixbits(uint32_t x) {   // bench.cu:152-155
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU;
    return x ^ (x >> 16);
}
__device__ __forceinline__ uint32_t u32hash(uint64_t seed, uint64_t g) { return uint32_t(splitmix64(seed ^ g) >> 32); }

constexpr uint64_t kSaltDup = 0x5bd1e9955bd1e995ull;
constexpr uint64_t kSaltSpecial = 0x00abcdef12345678ull;
constexpr uint64_t kSaltRun = 0x52554e5352554e53ull;
constexpr uint64_t kSaltRunBase = 0x4241534542415345ull;
constexpr uint64_t kSaltNormal = 0x4e4f524d4e4f524dull;

__device__ int32_t distinct16(uint32_t i) {
    switch (i & 15u) {
        case 0: return INT32_MIN; case 1: return -1000000000; case 2: return -65536; case 3: return -1000;
        case 4: return -7; case 5: return -1; case 6: return 0; case 7: return 1;
        case 8: return 2; case 9: return 3; case 10: return 42; case 11: return 1000;
        case 12: return 65536; case 13: return 1000000000; case 14: return INT32_MAX - 1; default: return INT32_MAX;
    }
}

__device__ int32_t dist_i32(int dist, int64_t n, int64_t g, uint64_t seed) {
    const int64_t p = g % n;
    switch (dist) {
        case CS_DIST_UNIFORM: return int32_t(u32hash(seed, uint64_t(g)));
        case CS_DIST_SORTED: return int32_t(p);
        case CS_DIST_REVERSED: return int32_t(n - 1 - p);
        case CS_DIST_CONSTANT: return 99;
        case CS_DIST_DISTINCT16: return distinct16(u32hash(seed, uint64_t(g)));
        case CS_DIST_PERTURBED: return int32_t(4 * p + int64_t(u32hash(seed, uint64_t(g)) % 17u) - 8);
        case CS_DIST_ORGAN_PIPE: return int32_t(p < n - 1 - p ? p : n - 1 - p);
        case CS_DIST_ALT_EXTREMES: return (p & 1) ? INT32_MAX : INT32_MIN;
        case CS_DIST_DUP50: {
            const int64_t half = (n + 1) / 2;
            const int64_t q = p < half ? p : int64_t(u32hash(seed ^ kSaltDup, uint64_t(g)) % uint64_t(half));
            return int32_t(u32hash(seed, uint64_t((g / n) * n + q)));
        }
        case CS_DIST_EXTREMES6: {
            switch (p % 6) {
                case 0: case 3: return INT32_MAX;
                case 1: return INT32_MIN;
                case 2: return 0;
                case 4: return -1;
                default: return 1;
            }
        }
        case CS_DIST_RANDOM1M: return int32_t(mixbits(uint32_t(g) + 1234u) % 1000001u);
        case CS_DIST_BINARY: return int32_t(p % 2);
        case CS_DIST_PRESORTED_RUNS: {
            const uint64_t block = uint64_t(g) / 4096u;
            if (u32hash(seed ^ kSaltRun, block) % 10u == 0)
                return int32_t(u32hash(seed ^ kSaltRunBase, block) >> 1) - (1 << 30) + int32_t(uint64_t(g) % 4096u);
            return int32_t(u32hash(seed, uint64_t(g)));
        }
        default: return 0;
    }
}

__device__ uint64_t dist_u64(int dist, int64_t n, int64_t g, uint64_t seed) {
    if (dist == CS_DIST_UNIFORM) return splitmix64(seed ^ uint64_t(g));
    if (dist == CS_DIST_PRESORTED_RUNS) {
        const uint64_t block = uint64_t(g) / 4096u;
        if (u32hash(seed ^ kSaltRun, block) % 10u == 0)
            return (splitmix64(seed ^ kSaltRunBase ^ block) >> 1) + uint64_t(g) % 4096u;
        return splitmix64(seed ^ uint64_t(g));
    }
    return uint64_t(int64_t(dist_i32(dist, n, g, seed)));
}

__device__ uint32_t dist_f32(int dist, int64_t n, int64_t g, uint64_t seed) {
    if (dist == CS_DIST_UNIFORM) return u32hash(seed, uint64_t(g));
    if (dist == CS_DIST_NORMAL_SPECIAL) {
        const uint32_t h = u32hash(seed ^ kSaltSpecial, uint64_t(g));
        if (h % 100u == 0) {
            switch ((h >> 8) & 7u) {
                case 0: return 0x7FC00000u; case 1: return 0xFFC00000u; case 2: return 0x7F800001u;
                case 3: return 0xFFFFFFFFu; case 4: return 0x80000000u; case 5: return 0x7F800000u;
                case 6: return 0xFF800000u; default: return 0x00000000u;
            }
        }
        int64_t sum = 0;
        for (uint64_t i = 0; i < 12; ++i) sum += int64_t(u32hash(seed ^ kSaltNormal, uint64_t(g) * 16u + i));
        sum -= int64_t(6) << 32;
        return __float_as_uint(float(double(sum) * 0x1p-32));
    }
    return __float_as_uint(float(dist_i32(dist, n, g, seed)));
}

// fp64: every bit pattern for UNIFORM; Irwin-Hall(12) - 6 in double plus 1 %
// specials (NaNs, infinities, -0, +0) for NORMAL_SPECIAL; the integer
// distributions converted to double otherwise.
__device__ uint64_t dist_f64(int dist, int64_t n, int64_t g, uint64_t seed) {
    if (dist == CS_DIST_UNIFORM) return splitmix64(seed ^ uint64_t(g));
    if (dist == CS_DIST_NORMAL_SPECIAL) {
        const uint32_t h = u32hash(seed ^ kSaltSpecial, uint64_t(g));
        if (h % 100u == 0) {
            switch ((h >> 8) & 7u) {
                case 0: return 0x7FF8000000000000ull; case 1: return 0xFFF8000000000000ull;
                case 2: return 0x7FF0000000000001ull; case 3: return 0xFFFFFFFFFFFFFFFFull;
                case 4: return 0x8000000000000000ull; case 5: return 0x7FF0000000000000ull;
                case 6: return 0xFFF0000000000000ull; default: return 0ull;
            }
        }
        int64_t sum = 0;
        for (uint64_t i = 0; i < 12; ++i) sum += int64_t(u32hash(seed ^ kSaltNormal, uint64_t(g) * 16u + i));
        sum -= int64_t(6) << 32;
        return uint64_t(__double_as_longlong(double(sum) * 0x1p-32));
    }
    return uint64_t(__double_as_longlong(double(dist_i32(dist, n, g, seed))));
}

__global__ void k_generate(int dtype, int dist, int64_t n, int64_t count, uint64_t seed, void* out) {
    for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < count; g += int64_t(gridDim.x) * blockDim.x) {
        switch (dtype) {
            case CS_I32: static_cast<int32_t*>(out)[g] = dist_i32(dist, n, g, seed); break;
            case CS_U32: static_cast<uint32_t*>(out)[g] = uint32_t(dist_i32(dist, n, g, seed)); break;
            case CS_F32: static_cast<uint32_t*>(out)[g] = dist_f32(dist, n, g, seed); break;
            case CS_U64: static_cast<uint64_t*>(out)[g] = dist_u64(dist, n, g, seed); break;
            case CS_I64: static_cast<int64_t*>(out)[g] = int64_t(dist_u64(dist, n, g, seed)); break;
            case CS_F64: static_cast<uint64_t*>(out)[g] = dist_f64(dist, n, g, seed); break;
            case CS_REC_I32_I32: {
                int32_t* r = static_cast<int32_t*>(out) + 2 * g;
                r[0] = dist_i32(dist, n, g, seed);
                r[1] = int32_t(g);
                break;
            }
            case CS_REC_I64_I32: {
                int64_t* r = static_cast<int64_t*>(out) + 2 * g;
                r[0] = int64_t(dist_u64(dist, n, g, seed));
                reinterpret_cast<int32_t*>(r + 1)[0] = int32_t(g);
                reinterpret_cast<int32_t*>(r + 1)[1] = 0;
                break;
            }
            default: break;
        }
    }
}

__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_copy1(const unsigned char* __restrict__ src, unsigned char* __restrict__ dst,
                                             size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

// Occupies its stream for `ticks` periods of the constant wall clock, then
// exits: a bounded wait (one wave, sleeping between clock reads).
__global__ void __launch_bounds__(64) k_stream_hold(uint64_t ticks) {
    const uint64_t t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(32);
}

}  // namespace

hipError_t stream_hold(double microseconds, hipStream_t s) {
    int dev = 0, khz = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    if (e != hipSuccess) return e;
    if (khz <= 0) return hipErrorInvalidValue;
    const uint64_t ticks = uint64_t(microseconds * double(khz) / 1000.0);
    hipLaunchKernelGGL(k_stream_hold, dim3(1), dim3(64), 0, s, ticks);
    return hipGetLastError();
}

hipError_t device_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    if ((((uintptr_t)dst | (uintptr_t)src | bytes) & 15) == 0) {
        const size_t n16 = bytes / 16;
        const size_t want = (n16 + 255) / 256;
        const unsigned blocks = unsigned(want < 8192 ? want : 8192);
        hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, static_cast<const uint4*>(src),
                           static_cast<uint4*>(dst), n16);
    } else {
        const size_t want = (bytes + 255) / 256;
        const unsigned blocks = unsigned(want < 8192 ? want : 8192);
        hipLaunchKernelGGL(k_copy1, dim3(blocks), dim3(256), 0, s, static_cast<const unsigned char*>(src),
                           static_cast<unsigned char*>(dst), bytes);
    }
    return hipGetLastError();
}

hipError_t generate(int dtype, int dist, int64_t n, int64_t count, uint64_t seed, void* out, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int64_t want = (count + 255) / 256;
    const unsigned blocks = unsigned(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(k_generate, dim3(blocks), dim3(256), 0, s, dtype, dist, n, count, seed, out);
    return hipGetLastError();
}

}  // namespace cs
