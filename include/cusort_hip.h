/* cusort-hip: C ABI of the MI355X (gfx950) sorting path.
 *
 * The reference (berenger-eu/cuda-sort) is a header-only CUDA C++ template
 * library with no C ABI; callers include cusort.hpp and launch its kernels.
 * This header is the boundary a foreign caller binds instead: plain pointers,
 * sizes and integer codes, stream-ordered on a caller-supplied hipStream_t.
 * Each entry point names the reference interface it replaces (file:line).
 *
 * Conventions
 *  - All device pointers are HIP device allocations owned by the caller.
 *  - Every call is asynchronous on `stream` (0 = the null stream) unless
 *    stated otherwise; errors in arguments are reported before anything is
 *    enqueued.
 *  - Return value: CS_OK (0), a negative CS_E_* code for argument/capacity
 *    errors, or a positive hipError_t from the HIP runtime.
 *  - Output equals host std::sort of each array (records compared by key;
 *    the order of equal keys is unspecified, as in the reference whose
 *    bitonic networks are unstable).
 */
#ifndef CUSORT_HIP_H
#define CUSORT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifndef HIP_INCLUDE_HIP_HIP_RUNTIME_API_H
typedef struct ihipStream_t* hipStream_t;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ element types
 * CS_I32          int32 key                     (reference `int`)
 * CS_REC_I32_I32  {int32 key; int32 payload}    (reference IntAndData,
 *                 intanddata.hpp:9-79: 8 B, compared by key only)
 * CS_F32          fp32 key ordered by IEEE-754 totalOrder
 *                 (-NaN < -inf < ... < -0 < +0 < ... < +inf < +NaN)
 * CS_U64          uint64 key
 * CS_REC_I64_I32  {int64 key; int32 payload; int32 pad} (16 B, key order)
 * CS_U32, CS_I64  uint32 / int64 keys
 * CS_F64          fp64 keys (raw bits), IEEE-754 totalOrder like CS_F32
 */
enum cs_dtype {
    CS_I32 = 0,
    CS_REC_I32_I32 = 1,
    CS_F32 = 2,
    CS_U64 = 3,
    CS_REC_I64_I32 = 4,
    CS_U32 = 5,
    CS_I64 = 6,
    CS_F64 = 7
};

/* Batched method names follow bench's --methods list (bench.cu:129-131). */
enum cs_method {
    CS_METHOD_AUTO = 0,          /* library's tuned choice                     */
    CS_BITONIC_WARP = 1,         /* Sort_1Warp_SM_le32 (cusort.hpp:574-606)     */
    CS_BITONIC_V1 = 2,           /* Sort_1Block_gt32_d_extsm (749-837)          */
    CS_BITONIC_V2 = 3,           /* ..._v2 (842-1056)                            */
    CS_BITONIC_V3 = 4,           /* ..._v3 (1061-1300)                           */
    CS_BITONIC_V4_XOR = 5,       /* cusort_experiments.hpp:180-228              */
    CS_BITONIC_V5_STATIC = 6,    /* cusort_experiments.hpp:159-233              */
    CS_BITONIC_V6_REGISTER = 7   /* cusort_v6.hpp:144-184                        */
};

/* Flags: the reference's compile-time switches, passed at run time.
 * TESTSORTED  skip the network when the array is already sorted (output
 *             unchanged, only time differs; cusort.hpp:394-458)
 * FULLSM      exchange every layer through LDS instead of registers
 * ASYNC       load key-only arrays of the LDS network straight into LDS
 *             (global_load_lds; memcpy_async analogue, cusort.hpp:808-837) */
#define CS_FLAG_TESTSORTED 1u
#define CS_FLAG_FULLSM 2u
#define CS_FLAG_ASYNC 4u

/* Status codes. */
#define CS_OK 0
#define CS_E_INVALID_VALUE (-1)     /* null pointer, negative size, bad enum  */
#define CS_E_UNSUPPORTED (-2)       /* (method, n, threads, flags, dtype) outside the domain */
#define CS_E_CAPACITY (-3)          /* length beyond the on-chip capacity      */
#define CS_E_NO_DEVICE (-4)         /* no HIP device / runtime failure at init */
#define CS_E_BUDGET (-5)            /* request exceeds an admission budget     */
#define CS_E_AGAIN (-6)             /* resource temporarily exhausted          */
#define CS_E_WORKSPACE (-7)         /* workspace too small / wrong type        */

const char* cs_strerror(int status);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int cs_version(void);
/* Size in bytes of one element of `dtype` in memory (0 for an unknown dtype). */
size_t cs_dtype_size(int dtype);
/* Map a bench method name ("bitonic_v6_register", "auto", ...) to cs_method;
 * returns -1 for an unknown name. */
int cs_method_from_name(const char* name);

/* ----------------------------------------------------------- batched sort
 * Sort `arrays` independent arrays of `n` elements each, packed back to back
 * (array a = base + a*n, BENCHMARKS.md:25-29), one workgroup (or a lane group
 * of one wave) per array.  in == out sorts in place (the reference's
 * behaviour); otherwise `out` receives the sorted arrays and `in` is left
 * untouched.
 *
 * Replaces the reference's batched launches
 *   benchmarkSortSmall / benchmarkSortBlock<1|2|3> / benchmarkSortV4Xor /
 *   benchmarkSortV5Static / benchmarkSortV6Register  (bench.cu:355-393)
 * launched as kernel<<<arrays, threads, W*sizeof(T)*(v1?1:2)>>>(data, n)
 * (bench.cu:442-552), with W = max(32, next_pow2(n)) (bench.cu:524).
 *
 * `threads` follows the reference's per-method domain (0 = method default):
 *   bitonic_warp        n <= 32, threads 32
 *   bitonic_v1, v2      threads == W, W <= 1024
 *   bitonic_v3          threads a multiple of 32 in [32, 1024] (above W the
 *                       extra waves idle; the reference's historical tests
 *                       run v3 so, test.cu:319-355)
 *   bitonic_v4_xor      threads a multiple of 32 in [32, min(W, 1024)]
 *   bitonic_v5_static   threads a power of two in [32, min(W, 1024)]
 *   bitonic_v6_register threads a power of two in [32, min(W, 1024)],
 *                       W / threads <= 16
 *   auto                threads ignored
 * Outside that domain the call returns CS_E_UNSUPPORTED (the reference
 * silently skipped unlaunchable configurations, bench.cu:425-440).
 * Capacity: n <= cs_batched_max_n(dtype) (16384 for 4-byte, 8192 for 8-byte,
 * 4096 for 16-byte elements: 160 KB of LDS per workgroup on gfx950, against the
 * reference's fixed 48 KB, cusort.hpp:231-237).
 */
int cs_batched_sort(int dtype, const void* in, void* out, int64_t arrays, int32_t n,
                    int method, int32_t threads, uint32_t flags, hipStream_t stream);
int32_t cs_batched_max_n(int dtype);
/* Key / payload pairs as two arrays (structure of arrays): keys of
 * key_dtype (CS_I32 or CS_I64) and int32 payloads, `arrays` arrays of n
 * packed back to back in each buffer, sorted by key.  No reference
 * counterpart as an API: it is the SoA form of the IntAndData record
 * (intanddata.hpp:9-79) and of the int64+int32 record, sorted by the same
 * network with the same tie rule, so the output equals sorting the records
 * {key[i], payload[i]} with cs_batched_sort (CS_REC_I32_I32 /
 * CS_REC_I64_I32, CS_METHOD_AUTO) and splitting them again.  Bytes per
 * element each way: 8 / 12 (the padded 16-byte record moves 16).
 * keys_in == keys_out and vals_in == vals_out sort in place.
 * n <= cs_batched_max_n of the record type (8192 / 4096). */
int cs_batched_sort_pairs(int key_dtype, const void* keys_in, const int32_t* vals_in, void* keys_out,
                          int32_t* vals_out, int64_t arrays, int32_t n, hipStream_t stream);
/* Resources of the kernel cs_batched_sort would launch: threads per
 * workgroup, items per thread, static LDS bytes, VGPRs, arrays per
 * workgroup (the reference reports the same through cudaFuncGetAttributes
 * and the occupancy API, bench.cu:425-440).  Any pointer may be NULL. */
int cs_batched_describe(int dtype, int32_t n, int method, int32_t threads, uint32_t flags,
                        int32_t* block_threads, int32_t* items_per_thread, int32_t* lds_bytes,
                        int32_t* vgprs, int32_t* arrays_per_block);

/* ------------------------------------------------------------ ragged batch
 * Sort every segment s = [offsets[s], offsets[s+1]) of a ragged batch given
 * as CSR offsets (int64, non-decreasing), each segment independently.
 * No reference counterpart as an API: the reference sorts ragged buckets
 * only inside its partition sort (SortPartitions, cusort.hpp:1802-1846).
 *
 * A plan is built once on the host offsets: segments are bucketed by
 * ceil(log2(length)), each bucket gets the tiling its tuned variant names,
 * work items are ordered largest bucket first, and one sort is one graph
 * (counter-reset kernel + one persistent kernel) launched on `stream`.
 * The plan keeps its graph; calls with other pointers update the graph's
 * kernel arguments.  One plan per host thread (plans are not thread-safe).
 * Longest segment: cs_segmented_max_len(dtype) (8192 for 4-byte keys,
 * 4096 otherwise). */
typedef struct cs_segplan cs_segplan;
#define CS_SEG_NO_GRAPH 1u         /* launch the kernels directly, no hipGraph */
int cs_segmented_plan_create(cs_segplan** plan, int dtype, const int64_t* h_offsets, int64_t segments,
                             uint32_t flags);
int cs_segmented_sort(cs_segplan* plan, const void* in, void* out, const int64_t* d_offsets, hipStream_t stream);
/* Key / payload pairs as two arrays (structure of arrays) through a plan
 * created for the record type (CS_REC_I32_I32 for int32 keys, CS_REC_I64_I32
 * for int64 keys) and int32 payloads.  No reference counterpart as an API:
 * the SoA form of the IntAndData record (intanddata.hpp:9-79).  The pairs of
 * [offsets[0], offsets[segments]) are packed into records in a stream-ordered
 * temporary, sorted by the plan and unpacked, so the output equals the record
 * sort of {key[i], payload[i]} split apart.  In place when keys_in ==
 * keys_out and vals_in == vals_out. */
int cs_segmented_sort_pairs(cs_segplan* plan, const void* keys_in, const int32_t* vals_in, void* keys_out,
                            int32_t* vals_out, const int64_t* d_offsets, hipStream_t stream);
int cs_segmented_plan_destroy(cs_segplan* plan);
/* Work items, persistent grid size, longest segment, and per bucket b < buckets
 * the chosen variant (0: 32-key register tile, 1: 16-key tile, 2: LDS network,
 * 3: cross-lane register engine for 4-byte keys without payload, W = 32..2048)
 * and its segment count.  Any pointer may be NULL. */
int cs_segmented_plan_info(const cs_segplan* plan, int32_t* work_items, int32_t* grid, int64_t* max_len,
                           int32_t* variants, int64_t* bucket_segments, int32_t buckets);
int32_t cs_segmented_max_len(int dtype);
/* Baseline: one lane group per segment, every segment in the tiling of the
 * longest one (the reference's SortPartitions pattern), one plain launch. */
int cs_segmented_sort_plain(int dtype, const void* in, void* out, const int64_t* d_offsets, int64_t segments,
                            int32_t max_len, hipStream_t stream);
/* Autotuner: time every variant of every bucket on full-length segments and
 * keep the fastest for plans created afterwards (per process, per dtype).
 * get/set read or override the table (-1 = untuned). */
int cs_segmented_tune(int dtype, int32_t trials, hipStream_t stream);
int cs_segmented_get_variant(int dtype, int32_t bucket, int32_t* variant);
int cs_segmented_set_variant(int dtype, int32_t bucket, int32_t variant);

/* ------------------------------------------------------------ large arrays
 * One array of n elements, sorted in place unless stated.
 *
 * cs_sort_merge replaces SortMerge<T, MaxValuesToSort, TestIfSorted, FullSM>
 * (cusort.hpp:2161-2300): chunks of 8192 (4-byte) / 4096 (8- and 16-byte)
 * elements are sorted by the register-tile kernel, then merge-path passes
 * double the sorted runs.  `tile` (MaxValuesToSort, power of two or 0) and
 * `merge_threads` are accepted for call compatibility; the chunk and merge
 * tile sizes are the library's.  ext_buf/ext_elems: optional scratch of at
 * least n elements; stored_in_right_place = 0 writes the result to ext_buf
 * (then required) and, as in the reference, uses the input as scratch.
 * Without scratch a stream-ordered temporary is allocated (hipMallocAsync).
 *
 * cs_sort_partition replaces Sort<T, TestIfSorted, FullSM, ASyncCpyBito,
 * MinNbParts, MaxNbParts> (cusort.hpp:2020-2023, 1869-2017): sample-sort
 * partition into up to 2*max_parts-1 buckets (equality buckets for repeated
 * splitters), one ragged sort of all buckets, merge passes for buckets longer
 * than a chunk.  Every step is laid out on the device, so the call only
 * enqueues work on the stream and never synchronises it (the reference
 * synchronises per recursion level, cusort.hpp:1977).  Without a workspace
 * its temporaries are stream-ordered allocations from a per-device pool that
 * keeps freed memory cached until the process ends.
 *
 * cs_sort_hybrid replaces HybridSort / HybridSortV3 (cusort.hpp:2656-2754):
 * merge below cs_hybrid_threshold(dtype) elements, partition above.
 *
 * Workspace (the reference's CuSortWorkspace, cusort.hpp:67-219): a values
 * buffer of `capacity` elements plus an arena holding every other temporary
 * of a merge or partition sort of up to `capacity` elements, and a pool of
 * n_streams worker streams with completion events.  The _ws entry points
 * and cs_sort_hybrid with a workspace of capacity >= n allocate no device
 * memory (cs_memory_stats counts the library's allocations); they fail with
 * CS_E_WORKSPACE when the workspace is too small or of another dtype.  A
 * workspace of capacity c holds every sort of n <= c elements with the
 * default part count or any max_parts; a min_parts above the default can
 * need more (CS_E_WORKSPACE).  Calls
 * that share a workspace are ordered through an event (each waits for the
 * previous one on its own stream); reserve and destroy free and allocate in
 * stream order behind them, without synchronising the host or the device.
 * One workspace per host thread (not thread-safe), as in the reference.     */
typedef struct cs_workspace cs_workspace;
typedef struct ihipEvent_t* cs_hip_event;
int cs_workspace_create(cs_workspace** ws, int dtype, size_t capacity, int32_t blocks_per_cu, int32_t n_streams);
int cs_workspace_reserve(cs_workspace* ws, size_t capacity);
int cs_workspace_capacity(const cs_workspace* ws, size_t* capacity);
/* Grows the arena to at least arena_bytes (stream ordered, like reserve). */
int cs_workspace_reserve_scratch(cs_workspace* ws, size_t arena_bytes);
/* setBlocksPerSM (cusort.hpp:170-188): bounds the partition path's persistent
 * grids to blocks_per_cu workgroups per CU. */
int cs_workspace_set_blocks_per_cu(cs_workspace* ws, int32_t blocks_per_cu);
/* valuesBuffer(), persistentBlocks(), nbSortStreams() (cusort.hpp:201-218)
 * and the arena; any pointer may be NULL. */
int cs_workspace_info(const cs_workspace* ws, void** values, void** arena, size_t* arena_bytes,
                      int32_t* persistent_blocks, int32_t* n_streams);
/* sortStream(i), sortDone(i) (cusort.hpp:214-215).  done[0] is recorded on
 * the caller's stream at the end of every sort that uses the workspace; the
 * other events are never recorded (waiting on them is a no-op). */
int cs_workspace_stream(const cs_workspace* ws, int32_t index, hipStream_t* stream, cs_hip_event* done);
/* workerEventsRecorded() (cusort.hpp:216). */
int cs_workspace_events_recorded(const cs_workspace* ws, int32_t* recorded);
/* Sorts run through the workspace by path (merge, partition). */
int cs_workspace_calls(const cs_workspace* ws, int64_t* merge_calls, int64_t* partition_calls);
int cs_workspace_destroy(cs_workspace* ws);
/* Device allocations the library has made from its pool so far (count and
 * bytes, process wide). */
int cs_memory_stats(int64_t* pool_allocations, int64_t* pool_bytes);
int cs_sort_merge(int dtype, void* data, int64_t n, int32_t tile, int32_t merge_threads, void* ext_buf,
                  size_t ext_elems, int stored_in_right_place, hipStream_t stream);
int cs_sort_merge_ws(int dtype, void* data, int64_t n, cs_workspace* ws, hipStream_t stream);
int cs_sort_partition(int dtype, void* data, int64_t n, int32_t min_parts, int32_t max_parts, uint32_t flags,
                      hipStream_t stream);
/* HybridSortPartitionedV3 (cusort.hpp:2488-2653): the partition path with
 * the workspace's buffers. */
int cs_sort_partition_ws(int dtype, void* data, int64_t n, int32_t min_parts, int32_t max_parts, uint32_t flags,
                         cs_workspace* ws, hipStream_t stream);
int cs_sort_hybrid(int dtype, void* data, int64_t n, cs_workspace* ws, hipStream_t stream);
/* One array of n key / payload pairs as two arrays (key_dtype CS_I32 or
 * CS_I64, int32 payloads), sorted in place by key through the hybrid sort of
 * the matching record type (no reference counterpart as an API: the SoA form
 * of IntAndData, intanddata.hpp:9-79).  The pairs are packed into records in
 * a stream-ordered temporary, sorted (ws: NULL or a workspace of the record
 * type) and unpacked: the output equals the record sort split apart. */
int cs_sort_pairs(int key_dtype, void* keys, int32_t* vals, int64_t n, cs_workspace* ws, hipStream_t stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * it).  The in-place merge sort of cs_sort_merge, with its scratch (n
 * elements plus one int64 per merge tile) taken from a fixed per-device
 * array of cs_sort_merge_static_max_bytes() bytes (64 MiB: up to ~16M
 * 4-byte or ~8M 8-byte elements) that the library holds until the process
 * ends; larger n fails with CS_E_CAPACITY.  It allocates nothing.  Every call
 * on one device shares that device's array, so calls that overlap on one
 * device race; none of the array needs to start at zero. */
int cs_sort_merge_static(int dtype, void* data, int64_t n, hipStream_t stream);
size_t cs_sort_merge_static_max_bytes(void);
int64_t cs_hybrid_threshold(int dtype);

/* ----------------------------------------------------------------- serving
 * Continuous batching of sort requests on one GPU.  No reference
 * counterpart.  A request is a host array of n elements (in) and a host
 * output array (out, may equal in); it completes when out holds the sorted
 * array.  Requests are queued in order; whenever one of `slots` in-flight
 * slots is free the scheduler scans the pending requests in order and packs
 * into one ragged batch every request that keeps the batch within
 * max_batch_elems elements and max_batch_requests requests (0 picks 4M and
 * 4096) and within the largest free block of the device pool, sized to the
 * HBM admission budget `budget_bytes` (a request longer than one tile also
 * needs merge scratch of its own size).  A request that does not fit is
 * passed over, at most 16 times: after that nothing queued behind it is
 * admitted before it.  When nothing fits the scheduler waits for a batch to
 * complete or a new request to arrive.  Each batch is copied in, sorted (one
 * persistent launch plus merge passes for long requests) and copied out on
 * its slot's stream, so batches overlap on the device, and each completes
 * as soon as its own work finishes.  Submission fails with CS_E_BUDGET for a
 * request that could never fit.  Latency = arrival (the caller's timestamp
 * from cs_server_now(), or the submission time when arrival_s < 0) to the
 * moment the result is in `out`.  The server keeps every request until
 * released. */
typedef struct cs_server cs_server;
typedef struct cs_server_stats {
    int64_t completed;          /* requests completed successfully */
    int64_t rejected;           /* submissions refused (CS_E_BUDGET) */
    int64_t batches;            /* batches launched */
    int64_t keys;               /* elements sorted */
    double p50_s, p99_s, max_s, mean_s;   /* latency percentiles, seconds */
    double keys_per_s;          /* keys / (last completion - first arrival) */
    double mean_batch_requests; /* average requests per batch */
    int64_t peak_pool_bytes;    /* high-water mark of the device pool */
    int64_t max_inflight_batches;
    int64_t bypasses;           /* times a request was admitted ahead of an earlier one */
} cs_server_stats;
double cs_server_now(void);
int cs_server_create(cs_server** s, int dtype, size_t budget_bytes, int32_t slots, int64_t max_batch_elems,
                     int32_t max_batch_requests);
int cs_server_submit(cs_server* s, const void* host_in, void* host_out, int64_t n, double arrival_s,
                     uint64_t* ticket);
/* done = 1 once finished; status = the request's result code. */
int cs_server_poll(cs_server* s, uint64_t ticket, int32_t* done, int32_t* status);
/* Blocks until the request finishes (timeout_s < 0: no limit); returns its
 * status, or CS_E_AGAIN on time-out. */
int cs_server_wait(cs_server* s, uint64_t ticket, double timeout_s);
int cs_server_latency(cs_server* s, uint64_t ticket, double* arrival_s, double* done_s);
int cs_server_release(cs_server* s, uint64_t ticket);
int cs_server_drain(cs_server* s);
int cs_server_get_stats(cs_server* s, cs_server_stats* stats);
int cs_server_destroy(cs_server* s);
/* Failure injection for tests of the error path: the next batch launched
 * fails with `status` (nonzero) after its input copy has been queued; its
 * requests finish with that status once the copy has drained, and later
 * batches run normally. */
int cs_server_inject_launch_failure(cs_server* s, int32_t status);
/* Bytes of the device pool currently held by batches. */
int cs_server_pool_in_use(cs_server* s, int64_t* bytes);

/* Device-to-device copy of `bytes` bytes by a plain streaming kernel (16-byte
 * accesses when both pointers are 16-byte aligned).  The reference harness
 * restores every input with a device copy before each trial
 * (bench.cu:304-307, 335); the benchmark also uses this kernel, whose HBM
 * traffic is known exactly, to calibrate rocprofv3's FETCH_SIZE/WRITE_SIZE. */
int cs_device_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

/* Enqueues one small kernel that keeps `stream` busy for about
 * `microseconds` (0 .. 1e6; a bounded wait on the device's constant wall
 * clock, then it exits).  No counterpart in the reference: bench.cu brackets
 * its launches with events directly (bench.cu:337-348).  Work queued behind
 * it starts back to back once it ends, so events around a run of short
 * launches time the device, not the host's launch rate. */
int cs_stream_hold(double microseconds, hipStream_t stream);

/* Input distributions of cs_generate (numbering shared with the oracle). */
enum cs_dist {
    CS_DIST_UNIFORM = 0,        /* uniform over the full key range            */
    CS_DIST_SORTED = 1,         /* local position (bench.cu:164-167)          */
    CS_DIST_REVERSED = 2,       /* n-1-local (bench.cu:168)                   */
    CS_DIST_CONSTANT = 3,       /* 99 (bench.cu:182)                          */
    CS_DIST_DISTINCT16 = 4,     /* 16 distinct values incl. INT_MIN/INT_MAX   */
    CS_DIST_PERTURBED = 5,      /* 4*local + U{-8..8} (PERTURBED_2 spacing)   */
    CS_DIST_ORGAN_PIPE = 6,     /* min(local, n-1-local)                      */
    CS_DIST_ALT_EXTREMES = 7,   /* INT_MIN, INT_MAX, INT_MIN, ...             */
    CS_DIST_DUP50 = 8,          /* second half repeats hashed first-half keys */
    CS_DIST_EXTREMES6 = 9,      /* {INT_MAX,INT_MIN,0,INT_MAX,-1,1} cycle (bench.cu:184-187) */
    CS_DIST_RANDOM1M = 10,      /* mixBits(i+1234) % 1000001 (bench.cu:171)  */
    CS_DIST_BINARY = 11,        /* local % 2 (bench.cu:183)                   */
    CS_DIST_NORMAL_SPECIAL = 12,/* fp32 ~N(0,1) (Irwin-Hall 12) + 1% NaN/-0/inf */
    CS_DIST_PRESORTED_RUNS = 13 /* 10% of 4096-blocks ascending runs, rest uniform */
};

/* ------------------------------------------------------- input generators
 * Fill count elements with distribution `dist` (same numbering and formulas
 * as the host oracle: uniform, sorted, reversed, constant, 16 distinct,
 * perturbed, organ pipe, alternating extremes, 50% duplicates, the
 * reference's EXTREMES / RANDOM / BINARY (bench.cu:152-192), fp32
 * N(0,1)+specials, 10% presorted runs).  Element g is at local position
 * g % n of array g / n; records get payload g (bench.cu:189).              */
int cs_generate(int dtype, int dist, int64_t n, int64_t count, uint64_t seed, void* out,
                hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CUSORT_HIP_H */
