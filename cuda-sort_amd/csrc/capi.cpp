// C ABI of the batched sort, generators and common helpers: argument checks
// against the reference's (method, n, threads) domain, kernel choice, launch.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/cusort_hip.h"
#include "batched.hpp"
#include "gen.hpp"

namespace cs {

KernelInfo tile_kernel(int dtype, int logw, int loge) {
    const bool hi = logw >= kSplitLogW;
#define CS_TILE_CASE(code, name) \
    case code: return hi ? tile_kernel_##name##_hi(logw, loge) : tile_kernel_##name##_lo(logw, loge);
    switch (dtype) {
        CS_TILE_CASE(CS_I32, i32)
        CS_TILE_CASE(CS_U32, u32)
        CS_TILE_CASE(CS_F32, f32)
        CS_TILE_CASE(CS_U64, u64)
        CS_TILE_CASE(CS_F64, f64)
        CS_TILE_CASE(CS_I64, i64)
        CS_TILE_CASE(CS_REC_I32_I32, rec32)
        CS_TILE_CASE(CS_REC_I64_I32, rec64)
        default: return KernelInfo{};
    }
#undef CS_TILE_CASE
}

KernelInfo lds_kernel(int dtype) {
    switch (dtype) {
        case CS_I32: return lds_kernel_i32();
        case CS_U32: return lds_kernel_u32();
        case CS_F32: return lds_kernel_f32();
        case CS_U64: return lds_kernel_u64();
        case CS_F64: return lds_kernel_f64();
        case CS_I64: return lds_kernel_i64();
        case CS_REC_I32_I32: return lds_kernel_rec32();
        case CS_REC_I64_I32: return lds_kernel_rec64();
        default: return KernelInfo{};
    }
}

namespace {

bool valid_dtype(int d) { return d >= CS_I32 && d <= CS_F64; }

int ceil_log2(int64_t n) {
    int l = 0;
    while ((int64_t(1) << l) < n) ++l;
    return l;
}

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

struct Choice {
    KernelInfo k;
    int logw = 0;
    int threads = 0;   // LDS kernel block size
};

// Resolve (method, n, threads, flags) to a kernel, following the reference's
// per-method domain (bench.cu:523-552): v1..v6 only for n >= 32 with
// W = max(32, next_pow2(n)); bitonic_warp only for n <= 32 with 32 threads.
int choose(int dtype, int32_t n, int method, int32_t threads, uint32_t flags, Choice* c) {
    if (!valid_dtype(dtype)) return CS_E_INVALID_VALUE;
    if (method < CS_METHOD_AUTO || method > CS_BITONIC_V6_REGISTER) return CS_E_INVALID_VALUE;
    if (flags & ~(CS_FLAG_TESTSORTED | CS_FLAG_FULLSM | CS_FLAG_ASYNC)) return CS_E_INVALID_VALUE;
    if (n < 1 || threads < 0) return CS_E_INVALID_VALUE;
    if (n > cs_batched_max_n(dtype)) return CS_E_CAPACITY;
    const int logn = ceil_log2(n);
    const int logref = logn < 5 ? 5 : logn;          // W = max(32, pow2(n))
    const int wref = 1 << logref;
    const bool fullsm = (flags & CS_FLAG_FULLSM) != 0;
    const int tmax = wref < 1024 ? wref : 1024;

    auto use_lds = [&](int logw, int t) {
        c->k = lds_kernel(dtype);
        c->logw = logw;
        c->threads = t;
        return CS_OK;
    };
    auto use_tile = [&](int logw, int loge) {
        c->k = tile_kernel(dtype, logw, loge);
        c->logw = logw;
        return c->k.launch ? CS_OK : CS_E_UNSUPPORTED;
    };
    auto lds_default_threads = [&](int logw) {
        const int half = (1 << logw) / 2;
        return half < 32 ? 32 : (half > 256 ? 256 : half);
    };

    // FULLSM: every exchange through LDS.  The register tile's cross-lane
    // moves are LDS transposes, but the cross-lane engine that serves
    // key-only 4-byte arrays by default moves keys between lanes with DPP and
    // permlane instead: FULLSM takes the 16-per-lane register tile there
    // (int32), or the LDS network for element types without that tiling.
    auto use_fullsm_tile = [&](int logw) {
        const int loge = logw < 4 ? logw : 4;
        if (xlane_shape(dtype, logw, auto_loge(dtype, logw))) {
            const KernelInfo k = tile_kernel(dtype, logw, loge);
            if (k.launch && loge != auto_loge(dtype, logw)) return use_tile(logw, loge);
            return use_lds(logw < 5 ? 5 : logw, lds_default_threads(logw < 5 ? 5 : logw));
        }
        return use_tile(logw, auto_loge(dtype, logw));
    };
    // v3 / v4 / v5 with a power-of-two block of t threads over W elements:
    // the register tile of W / t elements per lane when one exists for the
    // element type (TESTSORTED is checked inside it), else the LDS network.
    auto use_block_tile = [&](int logw, int t) {
        if (is_pow2(t) && t <= (1 << logw)) {
            const int loge = logw - ceil_log2(t);
            if (loge >= 1 && loge <= 5) {
                const KernelInfo k = tile_kernel(dtype, logw, loge);
                const bool lane_moves = xlane_shape(dtype, logw, loge);   // not LDS-only
                if (k.launch && !(fullsm && lane_moves)) return use_tile(logw, loge);
            }
        }
        return use_lds(logw, t);
    };

    switch (method) {
        case CS_METHOD_AUTO: {
            if (fullsm) return use_fullsm_tile(logn);
            return use_tile(logn, auto_loge(dtype, logn));
        }
        case CS_BITONIC_WARP: {
            if (n > 32 || (threads != 0 && threads != 32)) return CS_E_UNSUPPORTED;
            if (fullsm) return use_fullsm_tile(logn);
            return use_tile(logn, auto_loge(dtype, logn));
        }
        case CS_BITONIC_V1:
        case CS_BITONIC_V2: {
            if (n < 32 || wref > 1024) return CS_E_UNSUPPORTED;
            if (threads != 0 && threads != wref) return CS_E_UNSUPPORTED;
            return use_lds(logref, wref);
        }
        case CS_BITONIC_V3:
        case CS_BITONIC_V4_XOR: {
            if (n < 32) return CS_E_UNSUPPORTED;
            const int t = threads ? threads : (tmax < 256 ? tmax : 256);
            // v3 takes any multiple of 32 up to the block limit, also above W
            // (the reference's historical tests run it with 32..512 threads
            // at every size, test.cu:319-355); v4 keeps bench.cu's <= W.
            const int tcap = method == CS_BITONIC_V3 ? 1024 : tmax;
            if (t % 32 != 0 || t < 32 || t > tcap) return CS_E_UNSUPPORTED;
            return use_block_tile(logref, t);
        }
        case CS_BITONIC_V5_STATIC: {
            if (n < 32) return CS_E_UNSUPPORTED;
            const int t = threads ? threads : (tmax < 256 ? tmax : 256);
            if (!is_pow2(t) || t < 32 || t > tmax) return CS_E_UNSUPPORTED;
            return use_block_tile(logref, t);
        }
        case CS_BITONIC_V6_REGISTER: {
            if (n < 32) return CS_E_UNSUPPORTED;
            int t = threads;
            if (t == 0) t = wref / 16 < 32 ? 32 : wref / 16;
            if (!is_pow2(t) || t < 32 || t > tmax || wref / t > 16) return CS_E_UNSUPPORTED;
            const int loge = logref - ceil_log2(t);
            const KernelInfo k = tile_kernel(dtype, logref, loge);
            if (k.launch) return use_tile(logref, loge);   // both flags are handled inside the tile
            return use_lds(logref, t);   // dtypes without that tiling: same output
        }
        default: return CS_E_INVALID_VALUE;
    }
}

bool overlaps_partially(const void* a, const void* b, size_t bytes) {
    const char* x = static_cast<const char*>(a);
    const char* y = static_cast<const char*>(b);
    if (x == y) return false;
    return x < y + bytes && y < x + bytes;
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" {

const char* cs_strerror(int status) {
    switch (status) {
        case CS_OK: return "success";
        case CS_E_INVALID_VALUE: return "invalid argument";
        case CS_E_UNSUPPORTED: return "configuration outside the method's domain";
        case CS_E_CAPACITY: return "length exceeds the on-chip capacity";
        case CS_E_NO_DEVICE: return "no usable HIP device";
        case CS_E_BUDGET: return "request exceeds the admission budget";
        case CS_E_AGAIN: return "resource temporarily exhausted";
        case CS_E_WORKSPACE: return "workspace too small or of another element type";
        default: break;
    }
    if (status > 0) return hipGetErrorString(static_cast<hipError_t>(status));
    return "unknown status";
}

int cs_version(void) { return 100; }

size_t cs_dtype_size(int dtype) {
    switch (dtype) {
        case CS_I32: case CS_U32: case CS_F32: return 4;
        case CS_U64: case CS_I64: case CS_F64: case CS_REC_I32_I32: return 8;
        case CS_REC_I64_I32: return 16;
        default: return 0;
    }
}

int cs_method_from_name(const char* name) {
    if (!name) return -1;
    static const char* names[] = {"auto", "bitonic_warp", "bitonic_v1", "bitonic_v2", "bitonic_v3",
                                  "bitonic_v4_xor", "bitonic_v5_static", "bitonic_v6_register"};
    for (int i = 0; i < 8; ++i)
        if (std::strcmp(name, names[i]) == 0) return i;
    return -1;
}

int32_t cs_batched_max_n(int dtype) {
    if (!valid_dtype(dtype)) return 0;
    return int32_t(1) << max_logw(dtype);
}

int cs_batched_sort(int dtype, const void* in, void* out, int64_t arrays, int32_t n, int method,
                    int32_t threads, uint32_t flags, hipStream_t stream) {
    if (arrays < 0 || n < 0) return CS_E_INVALID_VALUE;
    if (!valid_dtype(dtype)) return CS_E_INVALID_VALUE;
    if (arrays == 0 || n == 0) return CS_OK;
    if (!in || !out) return CS_E_INVALID_VALUE;
    const size_t bytes = size_t(arrays) * size_t(n) * cs_dtype_size(dtype);
    if (overlaps_partially(in, out, bytes)) return CS_E_INVALID_VALUE;
    if (arrays > (int64_t(1) << 40)) return CS_E_INVALID_VALUE;
    Choice c;
    const int rc = choose(dtype, n, method, threads, flags, &c);
    if (rc != CS_OK) return rc;
    const hipError_t e = c.k.launch(in, out, arrays, n, c.logw, flags, c.threads, stream);
    return e == hipSuccess ? CS_OK : int(e);
}

int cs_batched_describe(int dtype, int32_t n, int method, int32_t threads, uint32_t flags, int32_t* block_threads,
                        int32_t* items_per_thread, int32_t* lds_bytes, int32_t* vgprs, int32_t* arrays_per_block) {
    Choice c;
    const int rc = choose(dtype, n, method, threads, flags, &c);
    if (rc != CS_OK) return rc;
    const bool lds = c.k.block == 0;
    const int w = 1 << c.logw;
    if (block_threads) *block_threads = lds ? c.threads : c.k.block;
    if (items_per_thread) *items_per_thread = lds ? (w + c.threads - 1) / c.threads : c.k.items;
    if (lds_bytes) {
        const int elem = int(cs_dtype_size(dtype) == 16 ? 12 : (dtype == CS_REC_I32_I32 ? 8 : cs_dtype_size(dtype)));
        *lds_bytes = lds ? w * elem : c.k.lds;
    }
    if (arrays_per_block) *arrays_per_block = c.k.apb;
    if (vgprs) {
        hipFuncAttributes attr;
        std::memset(&attr, 0, sizeof attr);
        if (hipFuncGetAttributes(&attr, c.k.kernel) == hipSuccess) *vgprs = attr.numRegs;
        else *vgprs = -1;
    }
    return CS_OK;
}

int cs_device_copy(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    if (bytes == 0) return CS_OK;
    if (!dst || !src) return CS_E_INVALID_VALUE;
    if (overlaps_partially(src, dst, bytes) || src == dst) return src == dst ? CS_OK : CS_E_INVALID_VALUE;
    const hipError_t e = cs::device_copy(dst, src, bytes, stream);
    return e == hipSuccess ? CS_OK : int(e);
}

int cs_stream_hold(double microseconds, hipStream_t stream) {
    if (!(microseconds >= 0.0 && microseconds <= 1e6)) return CS_E_INVALID_VALUE;
    const hipError_t e = cs::stream_hold(microseconds, stream);
    return e == hipSuccess ? CS_OK : int(e);
}

int cs_generate(int dtype, int dist, int64_t n, int64_t count, uint64_t seed, void* out, hipStream_t stream) {
    if (!valid_dtype(dtype) || dist < CS_DIST_UNIFORM || dist > CS_DIST_PRESORTED_RUNS) return CS_E_INVALID_VALUE;
    if (n < 1 || count < 0 || (count > 0 && !out)) return CS_E_INVALID_VALUE;
    const hipError_t e = cs::generate(dtype, dist, n, count, seed, out, stream);
    return e == hipSuccess ? CS_OK : int(e);
}

}  // extern "C"
