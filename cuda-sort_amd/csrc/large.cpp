// Large-array sorts: SortMerge / HybridSort / Sort entry points and the
// reusable workspace (the reference's CuSortWorkspace, cusort.hpp:67-219).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/cusort_hip.h"
#include "gen.hpp"
#include "large.hpp"
#include "scratch.hpp"

namespace cs {

LargeKernels large_kernels_i32();
LargeKernels large_kernels_u32();
LargeKernels large_kernels_f32();
LargeKernels large_kernels_u64();
LargeKernels large_kernels_i64();
LargeKernels large_kernels_f64();
LargeKernels large_kernels_rec32();
LargeKernels large_kernels_rec64();

std::atomic<int64_t> g_pool_allocs{0};
std::atomic<int64_t> g_pool_bytes{0};

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
    e = pool ? hipMallocFromPoolAsync(p, bytes, pool, s) : hipMallocAsync(p, bytes, s);
    if (e == hipSuccess) {
        g_pool_allocs.fetch_add(1, std::memory_order_relaxed);
        g_pool_bytes.fetch_add(int64_t(bytes), std::memory_order_relaxed);
    }
    return e;
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

size_t merge_scratch_bytes(int dtype, int64_t n) {
    const LargeKernels k = large_kernels_for(dtype);
    if (!k.chunk || n <= k.chunk) return 0;
    return scratch_round(size_t((n + k.tile - 1) / k.tile) * sizeof(int64_t));   // merge co-ranks
}

// Chunk sort + merge passes.  Result in `final` (== data for in place).
// `scratch` holds at least n elements and may be `data` itself for an
// out-of-place call (the input is then consumed, as in the reference).
int merge_sort(int dtype, void* data, int64_t n, void* final, void* scratch, Scratch& tmp) {
    const LargeKernels k = large_kernels_for(dtype);
    const hipStream_t stream = tmp.stream();
    if (n <= k.chunk) return cs_batched_sort(dtype, data, final, 1, int32_t(n), CS_METHOD_AUTO, 0, 0, stream);
    int passes = 0;
    for (int64_t run = k.chunk; run < n; run *= 2) ++passes;
    // Start co-rank of every merge tile (one int64 each).
    int64_t* cor = nullptr;
    const int rc = tmp.get(&cor, size_t((n + k.tile - 1) / k.tile));
    if (rc != CS_OK) return rc;
    // The chunk sort lands where an even number of ping-pong passes ends in `final`.
    void* a = (passes % 2 == 0) ? final : scratch;
    void* b = (a == final) ? scratch : final;
    hipError_t e = k.chunk_sort(data, a, n, stream);
    for (int64_t run = k.chunk; e == hipSuccess && run < n; run *= 2) {
        e = k.merge_pass(a, b, n, run, cor, stream);
        std::swap(a, b);
    }
    return e == hipSuccess ? CS_OK : int(e);
}

}  // namespace cs

using namespace cs;

// The reference's CuSortWorkspace (cusort.hpp:67-219): a values buffer of
// `capacity` elements, an arena for every other temporary of a sort of that
// size, and a pool of worker streams with completion events.  Its users and
// its own reallocations are ordered through `ready`: every call waits for it
// on the caller's stream and records it there afterwards, a reallocation
// frees and allocates in order behind it on the first worker stream, so
// nothing here synchronises the host or the device.
struct cs_workspace {
    int dtype = 0;
    void* values = nullptr;
    size_t capacity = 0;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    int blocks_per_cu = 4;
    int persistent_blocks = 1;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> done;
    hipEvent_t ready = nullptr;
    bool events_recorded = false;
    int64_t merge_calls = 0, partition_calls = 0;
};

namespace {

bool valid_dtype(int d) { return d >= CS_I32 && d <= CS_F64; }

int cu_count() {
    int dev = 0, cus = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1;
    return cus > 0 ? cus : 1;
}

// Arena bytes for sorts of up to `capacity` elements through either path,
// the partition path with its default parts or any smaller part bound: a
// bound of at most 256 parts takes one level with 64 samples per part and
// can need more scratch than the default two levels (both shapes are
// covered; fewer parts on two levels need less).
size_t arena_bound(int dtype, size_t capacity) {
    return std::max({merge_scratch_bytes(dtype, int64_t(capacity)),
                     partition_scratch_bytes(dtype, int64_t(capacity), 0, 0),
                     partition_scratch_bytes(dtype, int64_t(capacity), 0, 256)});
}

// Brackets one sort on `stream` that uses the workspace.
hipError_t ws_begin(cs_workspace* ws, hipStream_t stream) { return hipStreamWaitEvent(stream, ws->ready, 0); }
hipError_t ws_end(cs_workspace* ws, hipStream_t stream) {
    hipError_t e = hipEventRecord(ws->ready, stream);
    if (e == hipSuccess && !ws->done.empty()) e = hipEventRecord(ws->done[0], stream);
    if (e == hipSuccess) ws->events_recorded = true;
    return e;
}

}  // namespace

extern "C" {

int cs_workspace_create(cs_workspace** ws, int dtype, size_t capacity, int32_t blocks_per_cu, int32_t n_streams) {
    if (!ws || !valid_dtype(dtype) || blocks_per_cu < 1 || n_streams < 1) return CS_E_INVALID_VALUE;
    cs_workspace* w = new cs_workspace();
    w->dtype = dtype;
    w->blocks_per_cu = blocks_per_cu;
    w->persistent_blocks = blocks_per_cu * cu_count();
    hipError_t e = hipEventCreateWithFlags(&w->ready, hipEventDisableTiming);
    for (int i = 0; e == hipSuccess && i < n_streams; ++i) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) w->streams.push_back(st);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) w->done.push_back(ev);
    }
    int rc = e == hipSuccess ? CS_OK : int(e);
    if (rc == CS_OK && capacity) rc = cs_workspace_reserve(w, capacity);
    if (rc != CS_OK) {
        cs_workspace_destroy(w);
        *ws = nullptr;
        return rc;
    }
    *ws = w;
    return CS_OK;
}

namespace {
// Grows the values buffer and / or the arena, stream ordered on the first
// worker stream behind every earlier user of the workspace.
int ws_grow(cs_workspace* ws, size_t capacity, size_t arena_bytes) {
    const bool grow_values = capacity > ws->capacity;
    const bool grow_arena = arena_bytes > ws->arena_bytes;
    if (!grow_values && !grow_arena) return CS_OK;
    const hipStream_t s0 = ws->streams[0];
    hipError_t e = hipStreamWaitEvent(s0, ws->ready, 0);
    if (e == hipSuccess && grow_values && ws->values) e = hipFreeAsync(ws->values, s0);
    if (e == hipSuccess && grow_arena && ws->arena) e = hipFreeAsync(ws->arena, s0);
    if (grow_values) {
        ws->values = nullptr;
        ws->capacity = 0;
        if (e == hipSuccess) e = hipMallocAsync(&ws->values, capacity * cs_dtype_size(ws->dtype), s0);
        if (e == hipSuccess) ws->capacity = capacity;
    }
    if (grow_arena) {
        ws->arena = nullptr;
        ws->arena_bytes = 0;
        if (e == hipSuccess) e = hipMallocAsync(&ws->arena, arena_bytes, s0);
        if (e == hipSuccess) ws->arena_bytes = arena_bytes;
    }
    if (e == hipSuccess) e = hipEventRecord(ws->ready, s0);
    if (e != hipSuccess) (void)hipGetLastError();
    return e == hipSuccess ? CS_OK : int(e);
}
}  // namespace

int cs_workspace_reserve(cs_workspace* ws, size_t capacity) {
    if (!ws) return CS_E_INVALID_VALUE;
    if (capacity <= ws->capacity) return CS_OK;
    return ws_grow(ws, capacity, arena_bound(ws->dtype, capacity));
}

int cs_workspace_reserve_scratch(cs_workspace* ws, size_t arena_bytes) {
    if (!ws) return CS_E_INVALID_VALUE;
    return ws_grow(ws, ws->capacity, scratch_round(arena_bytes));
}

int cs_workspace_capacity(const cs_workspace* ws, size_t* capacity) {
    if (!ws || !capacity) return CS_E_INVALID_VALUE;
    *capacity = ws->capacity;
    return CS_OK;
}

int cs_workspace_set_blocks_per_cu(cs_workspace* ws, int32_t blocks_per_cu) {
    if (!ws || blocks_per_cu < 1) return CS_E_INVALID_VALUE;
    ws->blocks_per_cu = blocks_per_cu;
    ws->persistent_blocks = blocks_per_cu * cu_count();
    return CS_OK;
}

int cs_workspace_info(const cs_workspace* ws, void** values, void** arena, size_t* arena_bytes,
                      int32_t* persistent_blocks, int32_t* n_streams) {
    if (!ws) return CS_E_INVALID_VALUE;
    if (values) *values = ws->values;
    if (arena) *arena = ws->arena;
    if (arena_bytes) *arena_bytes = ws->arena_bytes;
    if (persistent_blocks) *persistent_blocks = ws->persistent_blocks;
    if (n_streams) *n_streams = int32_t(ws->streams.size());
    return CS_OK;
}

int cs_workspace_stream(const cs_workspace* ws, int32_t index, hipStream_t* stream, hipEvent_t* done) {
    if (!ws || index < 0 || size_t(index) >= ws->streams.size()) return CS_E_INVALID_VALUE;
    if (stream) *stream = ws->streams[size_t(index)];
    if (done) *done = ws->done[size_t(index)];
    return CS_OK;
}

int cs_workspace_events_recorded(const cs_workspace* ws, int32_t* recorded) {
    if (!ws || !recorded) return CS_E_INVALID_VALUE;
    *recorded = ws->events_recorded ? 1 : 0;
    return CS_OK;
}

int cs_workspace_calls(const cs_workspace* ws, int64_t* merge_calls, int64_t* partition_calls) {
    if (!ws) return CS_E_INVALID_VALUE;
    if (merge_calls) *merge_calls = ws->merge_calls;
    if (partition_calls) *partition_calls = ws->partition_calls;
    return CS_OK;
}

int cs_workspace_destroy(cs_workspace* ws) {
    if (!ws) return CS_OK;
    if (!ws->streams.empty()) {
        const hipStream_t s0 = ws->streams[0];
        if (ws->ready) (void)hipStreamWaitEvent(s0, ws->ready, 0);
        if (ws->values) (void)hipFreeAsync(ws->values, s0);
        if (ws->arena) (void)hipFreeAsync(ws->arena, s0);
    }
    for (hipEvent_t ev : ws->done) (void)hipEventDestroy(ev);
    for (hipStream_t st : ws->streams) (void)hipStreamDestroy(st);
    if (ws->ready) (void)hipEventDestroy(ws->ready);
    (void)hipGetLastError();
    delete ws;
    return CS_OK;
}

int cs_memory_stats(int64_t* pool_allocations, int64_t* pool_bytes) {
    if (pool_allocations) *pool_allocations = g_pool_allocs.load(std::memory_order_relaxed);
    if (pool_bytes) *pool_bytes = g_pool_bytes.load(std::memory_order_relaxed);
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
    Scratch tmp(stream);
    if (!stored_in_right_place) return merge_sort(dtype, data, n, final, data, tmp);
    if (ext_buf && ext_elems >= size_t(n)) return merge_sort(dtype, data, n, data, ext_buf, tmp);
    // No usable buffer: a stream-ordered temporary (cudaMallocAsync in the reference).
    void* values = nullptr;
    const int rc = tmp.get(reinterpret_cast<unsigned char**>(&values), size_t(n) * cs_dtype_size(dtype));
    if (rc != CS_OK) return rc;
    return merge_sort(dtype, data, n, data, values, tmp);
}

int cs_sort_merge_ws(int dtype, void* data, int64_t n, cs_workspace* ws, hipStream_t stream) {
    if (!ws || ws->dtype != dtype) return CS_E_WORKSPACE;
    if (!valid_dtype(dtype) || n < 0) return CS_E_INVALID_VALUE;
    if (n <= 1) return CS_OK;
    if (!data) return CS_E_INVALID_VALUE;
    if (ws->capacity < size_t(n) || merge_scratch_bytes(dtype, n) > ws->arena_bytes) return CS_E_WORKSPACE;
    hipError_t e = ws_begin(ws, stream);
    if (e != hipSuccess) return int(e);
    Scratch tmp(stream, ws->arena, ws->arena_bytes);
    const int rc = merge_sort(dtype, data, n, data, ws->values, tmp);
    ++ws->merge_calls;
    e = ws_end(ws, stream);
    return rc != CS_OK ? rc : (e == hipSuccess ? CS_OK : int(e));
}

int cs_sort_partition(int dtype, void* data, int64_t n, int32_t min_parts, int32_t max_parts, uint32_t flags,
                      hipStream_t stream) {
    if (!valid_dtype(dtype) || n < 0) return CS_E_INVALID_VALUE;
    if (min_parts < 0 || max_parts < 0 || (max_parts && min_parts > max_parts)) return CS_E_INVALID_VALUE;
    if (flags & ~(CS_FLAG_TESTSORTED | CS_FLAG_FULLSM | CS_FLAG_ASYNC)) return CS_E_INVALID_VALUE;
    if (n <= 1) return CS_OK;
    if (!data) return CS_E_INVALID_VALUE;
    Scratch tmp(stream);
    void* values = nullptr;
    const int rc = tmp.get(reinterpret_cast<unsigned char**>(&values), size_t(n) * cs_dtype_size(dtype));
    if (rc != CS_OK) return rc;
    return partition_sort(dtype, data, n, data, values, size_t(n), min_parts, max_parts, tmp);
}

int cs_sort_partition_ws(int dtype, void* data, int64_t n, int32_t min_parts, int32_t max_parts, uint32_t flags,
                         cs_workspace* ws, hipStream_t stream) {
    if (!ws || ws->dtype != dtype) return CS_E_WORKSPACE;
    if (!valid_dtype(dtype) || n < 0) return CS_E_INVALID_VALUE;
    if (min_parts < 0 || max_parts < 0 || (max_parts && min_parts > max_parts)) return CS_E_INVALID_VALUE;
    if (flags & ~(CS_FLAG_TESTSORTED | CS_FLAG_FULLSM | CS_FLAG_ASYNC)) return CS_E_INVALID_VALUE;
    if (n <= 1) return CS_OK;
    if (!data) return CS_E_INVALID_VALUE;
    if (ws->capacity < size_t(n) || partition_scratch_bytes(dtype, n, min_parts, max_parts) > ws->arena_bytes)
        return CS_E_WORKSPACE;
    hipError_t e = ws_begin(ws, stream);
    if (e != hipSuccess) return int(e);
    Scratch tmp(stream, ws->arena, ws->arena_bytes);
    tmp.grid_cap = ws->persistent_blocks;
    const int rc = partition_sort(dtype, data, n, data, ws->values, ws->capacity, min_parts, max_parts, tmp);
    ++ws->partition_calls;
    e = ws_end(ws, stream);
    return rc != CS_OK ? rc : (e == hipSuccess ? CS_OK : int(e));
}

int cs_sort_hybrid(int dtype, void* data, int64_t n, cs_workspace* ws, hipStream_t stream) {
    if (!valid_dtype(dtype) || n < 0) return CS_E_INVALID_VALUE;
    if (ws && ws->dtype != dtype) return CS_E_WORKSPACE;
    if (n <= 1) return CS_OK;
    if (!data) return CS_E_INVALID_VALUE;
    const bool partition = n >= cs_hybrid_threshold(dtype);
    if (ws && ws->capacity >= size_t(n))
        return partition ? cs_sort_partition_ws(dtype, data, n, 0, 0, 0, ws, stream)
                         : cs_sort_merge_ws(dtype, data, n, ws, stream);
    Scratch tmp(stream);
    void* values = nullptr;
    const int rc = tmp.get(reinterpret_cast<unsigned char**>(&values), size_t(n) * cs_dtype_size(dtype));
    if (rc != CS_OK) return rc;
    return partition ? partition_sort(dtype, data, n, data, values, size_t(n), 0, 0, tmp)
                     : merge_sort(dtype, data, n, data, values, tmp);
}

int64_t cs_hybrid_threshold(int dtype) {
    // Crossover measured on MI355X (uniform keys; merge / partition ms):
    //   int32                2^22 0.30 / 0.32   2^23 0.36 / 0.35   2^24 0.59 / 0.46
    //   uint64               2^22 0.37 / 0.49   2^23 0.61 / 0.55   2^24 1.08 / 0.77
    //   int32+int32 records  2^22 0.31 / 0.40   2^23 0.52 / 0.47   2^24 0.93 / 0.66
    //   int64+int32 records  2^20 0.34 / 0.44   2^21 0.60 / 0.52   2^22 1.10 / 0.62
    // The reference encodes its own crossover as CuSortTuning
    // (cusort.hpp:2476-2483: never for 4-byte keys, 2^27 for 8-byte ones).
    if (!valid_dtype(dtype)) return 0;
    const size_t size = cs_dtype_size(dtype);
    return int64_t(1) << (size >= 16 ? 21 : 23);
}

}  // extern "C"
