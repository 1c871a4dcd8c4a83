// Large-array sorts: SortMerge / HybridSort / Sort entry points and the
// reusable workspace (the reference's CuSortWorkspace, cusort.hpp:67-219).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "../../include/cusort_hip.h"
#include "gen.hpp"
#include "large.hpp"

namespace cs {

LargeKernels large_kernels_i32();
LargeKernels large_kernels_u32();
LargeKernels large_kernels_f32();
LargeKernels large_kernels_u64();
LargeKernels large_kernels_i64();
LargeKernels large_kernels_f64();
LargeKernels large_kernels_rec32();
LargeKernels large_kernels_rec64();

hipError_t pool_malloc_async(void** p, size_t bytes, hipStream_t s) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static hipMemPool_t pools[kMaxDevices] = {};
    static bool tried[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    hipMemPool_t pool = nullptr;
    if (dev >= 0 && dev < kMaxDevices) {
        std::lock_guard<std::mutex> lock(mu);
        if (!tried[dev]) {
            tried[dev] = true;
            hipMemPoolProps props = {};
            props.allocType = hipMemAllocationTypePinned;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            hipMemPool_t created = nullptr;
            if (hipMemPoolCreate(&created, &props) == hipSuccess) {
                uint64_t keep = UINT64_MAX;
                if (hipMemPoolSetAttribute(created, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess)
                    pools[dev] = created;
                else
                    (void)hipMemPoolDestroy(created);
            }
            (void)hipGetLastError();
        }
        pool = pools[dev];
    }
    // Without a pool of its own the device's default pool serves.
    return pool ? hipMallocFromPoolAsync(p, bytes, pool, s) : hipMallocAsync(p, bytes, s);
}

LargeKernels large_kernels_for(int dtype) {
    switch (dtype) {
        case CS_I32: return large_kernels_i32();
        case CS_U32: return large_kernels_u32();
        case CS_F32: return large_kernels_f32();
        case CS_U64: return large_kernels_u64();
        case CS_I64: return large_kernels_i64();
        case CS_F64: return large_kernels_f64();
        case CS_REC_I32_I32: return large_kernels_rec32();
        case CS_REC_I64_I32: return large_kernels_rec64();
        default: return LargeKernels{};
    }
}

int partition_sort(int dtype, void* data, int64_t n, void* out, void* scratch, size_t scratch_elems, int min_parts,
                   int max_parts, hipStream_t stream);

}  // namespace cs

using namespace cs;

struct cs_workspace {
    int dtype = 0;
    void* values = nullptr;
    size_t capacity = 0;
    int blocks_per_cu = 4;
    int streams = 8;
};

namespace {

bool valid_dtype(int d) { return d >= CS_I32 && d <= CS_F64; }

// Chunk sort + merge passes.  Result in `final` (== data for in place).
// `scratch` holds at least n elements and may be `data` itself for an
// out-of-place call (the input is then consumed, as in the reference).
int merge_sort(int dtype, void* data, int64_t n, void* final, void* scratch, hipStream_t stream) {
    const LargeKernels k = large_kernels_for(dtype);
    if (n <= k.chunk) {
        const int rc = cs_batched_sort(dtype, data, final, 1, int32_t(n), CS_METHOD_AUTO, 0, 0, stream);
        return rc;
    }
    int passes = 0;
    for (int64_t run = k.chunk; run < n; run *= 2) ++passes;
    // Start co-rank of every merge tile (one int64 each), stream ordered.
    const int64_t tiles = (n + k.tile - 1) / k.tile;
    int64_t* cor = nullptr;
    hipError_t e = pool_malloc_async(reinterpret_cast<void**>(&cor), size_t(tiles) * sizeof(int64_t), stream);
    if (e != hipSuccess) return int(e);
    // The chunk sort lands where an even number of ping-pong passes ends in `final`.
    void* a = (passes % 2 == 0) ? final : scratch;
    void* b = (a == final) ? scratch : final;
    e = k.chunk_sort(data, a, n, stream);
    for (int64_t run = k.chunk; e == hipSuccess && run < n; run *= 2) {
        e = k.merge_pass(a, b, n, run, cor, stream);
        std::swap(a, b);
    }
    const hipError_t fe = hipFreeAsync(cor, stream);
    if (e == hipSuccess) e = fe;
    return e == hipSuccess ? CS_OK : int(e);
}

}  // namespace

extern "C" {

int cs_workspace_create(cs_workspace** ws, int dtype, size_t capacity, int32_t blocks_per_cu, int32_t n_streams) {
    if (!ws || !valid_dtype(dtype) || blocks_per_cu < 1 || n_streams < 1) return CS_E_INVALID_VALUE;
    cs_workspace* w = new cs_workspace();
    w->dtype = dtype;
    w->blocks_per_cu = blocks_per_cu;
    w->streams = n_streams;
    *ws = w;
    if (capacity) {
        const int rc = cs_workspace_reserve(w, capacity);
        if (rc != CS_OK) { delete w; *ws = nullptr; return rc; }
    }
    return CS_OK;
}

int cs_workspace_reserve(cs_workspace* ws, size_t capacity) {
    if (!ws) return CS_E_INVALID_VALUE;
    if (capacity <= ws->capacity) return CS_OK;
    if (ws->values) {
        hipError_t e = hipDeviceSynchronize();   // the old buffer may still be in use by queued work
        if (e == hipSuccess) e = hipFree(ws->values);
        ws->values = nullptr;
        ws->capacity = 0;
        if (e != hipSuccess) return int(e);
    }
    const hipError_t e = hipMalloc(&ws->values, capacity * cs_dtype_size(ws->dtype));
    if (e != hipSuccess) return int(e);
    ws->capacity = capacity;
    return CS_OK;
}

int cs_workspace_capacity(const cs_workspace* ws, size_t* capacity) {
    if (!ws || !capacity) return CS_E_INVALID_VALUE;
    *capacity = ws->capacity;
    return CS_OK;
}

int cs_workspace_destroy(cs_workspace* ws) {
    if (!ws) return CS_OK;
    if (ws->values) {
        (void)hipDeviceSynchronize();
        (void)hipFree(ws->values);
    }
    delete ws;
    return CS_OK;
}

int cs_sort_merge(int dtype, void* data, int64_t n, int32_t tile, int32_t merge_threads, void* ext_buf,
                  size_t ext_elems, int stored_in_right_place, hipStream_t stream) {
    if (!valid_dtype(dtype) || n < 0) return CS_E_INVALID_VALUE;
    if (tile != 0 && (tile < 32 || (tile & (tile - 1)) != 0)) return CS_E_INVALID_VALUE;
    if (merge_threads < 0 || merge_threads > 1024) return CS_E_INVALID_VALUE;
    if (!stored_in_right_place && (!ext_buf || ext_elems < size_t(n))) return CS_E_INVALID_VALUE;
    if (n == 0) return CS_OK;
    if (!data) return CS_E_INVALID_VALUE;
    void* final = stored_in_right_place ? data : ext_buf;
    if (n == 1) {
        if (final == data) return CS_OK;
        return cs_device_copy(final, data, cs_dtype_size(dtype), stream);
    }
    if (!stored_in_right_place) return merge_sort(dtype, data, n, final, data, stream);
    if (ext_buf && ext_elems >= size_t(n)) return merge_sort(dtype, data, n, data, ext_buf, stream);
    // No usable buffer: a stream-ordered temporary (cudaMallocAsync in the reference).
    void* tmp = nullptr;
    hipError_t e = pool_malloc_async(&tmp, size_t(n) * cs_dtype_size(dtype), stream);
    if (e != hipSuccess) return int(e);
    const int rc = merge_sort(dtype, data, n, data, tmp, stream);
    e = hipFreeAsync(tmp, stream);
    return rc != CS_OK ? rc : (e == hipSuccess ? CS_OK : int(e));
}

int cs_sort_merge_ws(int dtype, void* data, int64_t n, cs_workspace* ws, hipStream_t stream) {
    if (!ws || ws->dtype != dtype) return CS_E_WORKSPACE;
    if (n < 0) return CS_E_INVALID_VALUE;
    if (n > 1 && ws->capacity < size_t(n)) return CS_E_WORKSPACE;
    return cs_sort_merge(dtype, data, n, 0, 0, ws->values, ws->capacity, 1, stream);
}

int cs_sort_partition(int dtype, void* data, int64_t n, int32_t min_parts, int32_t max_parts, uint32_t flags,
                      hipStream_t stream) {
    if (!valid_dtype(dtype) || n < 0) return CS_E_INVALID_VALUE;
    if (min_parts < 0 || max_parts < 0 || (max_parts && min_parts > max_parts)) return CS_E_INVALID_VALUE;
    if (flags & ~(CS_FLAG_TESTSORTED | CS_FLAG_FULLSM | CS_FLAG_ASYNC)) return CS_E_INVALID_VALUE;
    if (n <= 1) return CS_OK;
    if (!data) return CS_E_INVALID_VALUE;
    void* tmp = nullptr;
    hipError_t e = pool_malloc_async(&tmp, size_t(n) * cs_dtype_size(dtype), stream);
    if (e != hipSuccess) return int(e);
    const int rc = partition_sort(dtype, data, n, data, tmp, size_t(n), min_parts, max_parts, stream);
    e = hipFreeAsync(tmp, stream);
    return rc != CS_OK ? rc : (e == hipSuccess ? CS_OK : int(e));
}

int cs_sort_hybrid(int dtype, void* data, int64_t n, cs_workspace* ws, hipStream_t stream) {
    if (!valid_dtype(dtype) || n < 0) return CS_E_INVALID_VALUE;
    if (ws && ws->dtype != dtype) return CS_E_WORKSPACE;
    if (n <= 1) return CS_OK;
    if (!data) return CS_E_INVALID_VALUE;
    const int64_t threshold = cs_hybrid_threshold(dtype);
    void* scratch = nullptr;
    bool owned = false;
    if (ws && ws->capacity >= size_t(n)) {
        scratch = ws->values;
    } else {
        const hipError_t e = pool_malloc_async(&scratch, size_t(n) * cs_dtype_size(dtype), stream);
        if (e != hipSuccess) return int(e);
        owned = true;
    }
    int rc;
    if (n < threshold) rc = merge_sort(dtype, data, n, data, scratch, stream);
    else rc = partition_sort(dtype, data, n, data, scratch, size_t(n), 0, 0, stream);
    if (owned) {
        const hipError_t e = hipFreeAsync(scratch, stream);
        if (rc == CS_OK && e != hipSuccess) rc = int(e);
    }
    return rc;
}

int64_t cs_hybrid_threshold(int dtype) {
    // Crossover measured on MI355X (uniform keys; merge / partition ms):
    //   int32                2^24 0.56 / 0.83   2^26 2.54 / 1.70   2^28 10.3 / 5.23
    //   uint64               2^22 0.39 / 0.87   2^24 1.23 / 1.15   2^26 5.35 / 2.78
    //   int32+int32 records  2^24 0.95 / 1.00   2^25 2.16 / 1.44
    //   int64+int32 records  2^21 0.59 / 0.80   2^22 1.09 / 0.92
    // The reference encodes its own crossover as CuSortTuning
    // (cusort.hpp:2476-2483: never for 4-byte keys, 2^27 for 8-byte ones).
    if (!valid_dtype(dtype)) return 0;
    const size_t size = cs_dtype_size(dtype);
    return int64_t(1) << (size >= 16 ? 22 : (size >= 8 ? 24 : 25));
}

}  // extern "C"
