// Ragged-batch plans: bucketing, work items, the hipGraph of one sort, and
// the per-bucket tuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/cusort_hip.h"
#include "gen.hpp"
#include "ragged.hpp"

namespace cs {
int pair_record_dtype(int key_dtype);
hipError_t pairs_pack(int key_dtype, const void* keys, const int32_t* vals, void* rec, int64_t n, hipStream_t s);
hipError_t pairs_unpack(int key_dtype, const void* rec, void* keys, int32_t* vals, int64_t n, hipStream_t s);
}  // namespace cs

namespace cs {

__global__ void k_seg_reset(int* counters) {
    if (blockIdx.x == 0 && threadIdx.x < 2) counters[threadIdx.x] = 0;
}

SegKernels seg_kernels_i32();
SegKernels seg_kernels_u32();
SegKernels seg_kernels_f32();
SegKernels seg_kernels_u64();
SegKernels seg_kernels_i64();
SegKernels seg_kernels_f64();
SegKernels seg_kernels_rec32();
SegKernels seg_kernels_rec64();

SegKernels seg_kernels_for(int dtype) {
    switch (dtype) {
        case CS_I32: return seg_kernels_i32();
        case CS_U32: return seg_kernels_u32();
        case CS_F32: return seg_kernels_f32();
        case CS_U64: return seg_kernels_u64();
        case CS_I64: return seg_kernels_i64();
        case CS_F64: return seg_kernels_f64();
        case CS_REC_I32_I32: return seg_kernels_rec32();
        case CS_REC_I64_I32: return seg_kernels_rec64();
        default: return SegKernels{};
    }
}

namespace {

constexpr int kMaxBuckets = 16;

// Tuned variant per (dtype, bucket); -1 = not tuned (use variant 0).
std::mutex g_tune_mu;
int g_variant[8][kMaxBuckets];
bool g_variant_init = false;

int tuned_variant(int dtype, int b) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!g_variant_init) {
        for (auto& row : g_variant)
            for (int& v : row) v = -1;
        g_variant_init = true;
    }
    return g_variant[dtype][b];
}

void set_tuned_variant(int dtype, int b, int v) {
    tuned_variant(dtype, b);   // initialise
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_variant[dtype][b] = v;
}

int bucket_of(int64_t len) {   // ceil(log2(len)), 0 for len <= 1
    return len > 1 ? 64 - __builtin_clzll(uint64_t(len - 1)) : 0;
}

}  // namespace
}  // namespace cs

using namespace cs;

struct cs_segplan {
    int dtype = 0;
    int64_t segments = 0;
    int64_t base = 0;    // first offset
    int64_t total = 0;
    int64_t max_len = 0;
    int nitems = 0;
    SegShape shape;
    int variant[kMaxBuckets] = {};
    int64_t bucket_segments[kMaxBuckets] = {};
    SegKernels k;
    int* d_list = nullptr;
    SegItem* d_items = nullptr;
    int* d_counter = nullptr;          // two item counters
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t reset_node = nullptr;
    hipGraphNode_t sort_node[2] = {nullptr, nullptr};
    // arguments baked into the graph's nodes
    const void* g_in = nullptr;
    void* g_out = nullptr;
    const int64_t* g_off = nullptr;
    void* args[2][9] = {};
    void* reset_args[1] = {};
    const void* a_in = nullptr;
    void* a_out = nullptr;
    const int64_t* a_off = nullptr;
    const int* a_list = nullptr;
    const SegItem* a_items = nullptr;
    int a_begin[2] = {0, 0};
    int a_end[2] = {0, 0};
    int* a_counter[2] = {nullptr, nullptr};
    int* a_reset = nullptr;
    const int* a_range = nullptr;   // item ranges come from the host
};

namespace {

void free_plan(cs_segplan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_list) (void)hipFree(p->d_list);
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_counter) (void)hipFree(p->d_counter);
    delete p;
}

hipKernelNodeParams sort_node_params(cs_segplan* p, int w) {
    hipKernelNodeParams kp{};
    kp.func = const_cast<void*>(p->k.persistent_fn[w]);
    kp.gridDim = dim3(unsigned(p->shape.grid[w]));
    kp.blockDim = dim3(kSegBlock);
    kp.sharedMemBytes = 0;
    kp.kernelParams = p->args[w];
    kp.extra = nullptr;
    return kp;
}

// The graph: counter reset (a kernel node, not a memset node), then the
// two persistent kernels as independent nodes, free to run concurrently.
hipError_t build_graph(cs_segplan* p, const void* in, void* out, const int64_t* off) {
    if (p->exec) { (void)hipGraphExecDestroy(p->exec); p->exec = nullptr; }
    if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
    p->sort_node[0] = p->sort_node[1] = nullptr;
    hipError_t e = hipGraphCreate(&p->graph, 0);
    if (e != hipSuccess) return e;
    p->a_reset = p->d_counter;
    p->reset_args[0] = &p->a_reset;
    hipKernelNodeParams rp{};
    rp.func = reinterpret_cast<void*>(&k_seg_reset);
    rp.gridDim = dim3(1);
    rp.blockDim = dim3(64);
    rp.sharedMemBytes = 0;
    rp.kernelParams = p->reset_args;
    rp.extra = nullptr;
    e = hipGraphAddKernelNode(&p->reset_node, p->graph, nullptr, 0, &rp);
    if (e != hipSuccess) return e;
    p->a_in = in; p->a_out = out; p->a_off = off;
    p->a_list = p->d_list; p->a_items = p->d_items;
    for (int w = 0; w < 2; ++w) {
        p->a_begin[w] = w == kSegMulti ? 0 : p->shape.split;
        p->a_end[w] = w == kSegMulti ? p->shape.split : p->nitems;
        p->a_counter[w] = p->d_counter + w;
        void** a = p->args[w];
        a[0] = &p->a_in; a[1] = &p->a_out; a[2] = &p->a_off; a[3] = &p->a_list; a[4] = &p->a_items;
        a[5] = &p->a_begin[w]; a[6] = &p->a_end[w]; a[7] = &p->a_counter[w]; a[8] = &p->a_range;
        if (p->a_end[w] <= p->a_begin[w]) continue;
        const hipKernelNodeParams kp = sort_node_params(p, w);
        e = hipGraphAddKernelNode(&p->sort_node[w], p->graph, &p->reset_node, 1, &kp);
        if (e != hipSuccess) return e;
    }
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return e;
    p->g_in = in; p->g_out = out; p->g_off = off;
    return hipSuccess;
}

hipError_t update_graph(cs_segplan* p, const void* in, void* out, const int64_t* off) {
    p->a_in = in; p->a_out = out; p->a_off = off;
    hipError_t e = hipSuccess;
    for (int w = 0; w < 2 && e == hipSuccess; ++w) {
        if (!p->sort_node[w]) continue;
        const hipKernelNodeParams kp = sort_node_params(p, w);
        e = hipGraphExecKernelNodeSetParams(p->exec, p->sort_node[w], &kp);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return build_graph(p, in, out, off);   // fall back to a fresh instantiation
    }
    p->g_in = in; p->g_out = out; p->g_off = off;
    return hipSuccess;
}

int make_plan_impl(cs_segplan** out, int dtype, const int64_t* h_offsets, int64_t segments, uint32_t flags,
                   const int* force_variant) {
    if (!out || segments < 0 || (segments > 0 && !h_offsets)) return CS_E_INVALID_VALUE;
    if (dtype < CS_I32 || dtype > CS_F64) return CS_E_INVALID_VALUE;
    if (flags & ~(CS_SEG_NO_GRAPH)) return CS_E_INVALID_VALUE;
    if (segments > (int64_t(1) << 31) - 1) return CS_E_INVALID_VALUE;
    cs_segplan* p = new cs_segplan();
    p->dtype = dtype;
    p->segments = segments;
    p->k = seg_kernels_for(dtype);
    p->use_graph = (flags & CS_SEG_NO_GRAPH) == 0;
    std::vector<int> list;
    std::vector<SegItem> items;
    const int rc0 = seg_build_items(dtype, h_offsets, segments, force_variant, list, items, &p->max_len, p->variant,
                                    p->bucket_segments);
    if (rc0 != CS_OK) { delete p; return rc0; }
    p->base = segments ? h_offsets[0] : 0;
    p->total = segments ? h_offsets[segments] - h_offsets[0] : 0;
    p->nitems = int(items.size());
    p->shape = seg_shape(dtype, items);
    if (p->shape.grid[0] <= 0 || p->shape.grid[1] <= 0) { free_plan(p); return CS_E_NO_DEVICE; }
    hipError_t e = hipMalloc(&p->d_list, std::max<size_t>(list.size(), 1) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&p->d_items, std::max<size_t>(items.size(), 1) * sizeof(SegItem));
    if (e == hipSuccess) e = hipMalloc(&p->d_counter, 2 * sizeof(int));
    if (e == hipSuccess && !list.empty())
        e = hipMemcpy(p->d_list, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess && !items.empty())
        e = hipMemcpy(p->d_items, items.data(), items.size() * sizeof(SegItem), hipMemcpyHostToDevice);
    if (e != hipSuccess) { free_plan(p); return int(e); }
    *out = p;
    return CS_OK;
}

int run_plan(cs_segplan* p, const void* in, void* out, const int64_t* d_offsets, hipStream_t stream) {
    if (!p) return CS_E_INVALID_VALUE;
    if (p->nitems == 0) return CS_OK;
    if (!in || !out || !d_offsets) return CS_E_INVALID_VALUE;
    hipError_t e;
    if (p->use_graph) {
        if (!p->exec) e = build_graph(p, in, out, d_offsets);
        else if (p->g_in != in || p->g_out != out || p->g_off != d_offsets) e = update_graph(p, in, out, d_offsets);
        else e = hipSuccess;
        if (e == hipSuccess) e = hipGraphLaunch(p->exec, stream);
    } else {
        e = seg_launch_all(p->dtype, p->shape, in, out, d_offsets, p->d_list, p->d_items, p->nitems, p->d_counter,
                           stream);
    }
    return e == hipSuccess ? CS_OK : int(e);
}

}  // namespace

namespace cs {

// Variant of a bucket the tuner has not set: the cross-lane engine where it
// applies (4-byte keys without payload, W = 32..2048; measured on 2^28 int32
// partition 3.32 -> 3.29 ms and the 100,000-segment ragged batch 0.539 ->
// 0.537 ms), else the 32-key register tile.
static int default_variant(const SegKernels& k, int b) { return k.variant_ok(b, 3) ? 3 : 0; }

SegClasses seg_classes(int dtype) {
    const SegKernels k = seg_kernels_for(dtype);
    SegClasses c;
    c.maxb = k.max_logw;
    c.small_max = k.small_max;
    for (int b = 0; b <= c.maxb; ++b) {
        int v = tuned_variant(dtype, b);
        if (v < 0 || !k.variant_ok(b, v)) v = default_variant(k, b);
        if (!k.variant_ok(b, v)) v = 2;
        c.variant[b] = v;
        c.spi[b] = k.segments_per_item(b, v);
    }
    return c;
}

int seg_build_items(int dtype, const int64_t* h_off, int64_t segs, const int* force_variant, std::vector<int>& list,
                    std::vector<SegItem>& items, int64_t* max_len, int* variants, int64_t* bucket_segments) {
    const SegKernels k = seg_kernels_for(dtype);
    const int maxb = k.max_logw;
    if (segs > kSegIndex) return CS_E_CAPACITY;   // list entries keep two flag bits
    // counting sort of the segments by bucket, largest bucket first
    thread_local std::vector<unsigned char> bkt;   // kept mapped between calls
    bkt.resize(static_cast<size_t>(segs));
    int64_t count[kMaxBuckets] = {};
    int64_t mx = 0;
    for (int64_t s = 0; s < segs; ++s) {
        const int64_t len = h_off[s + 1] - h_off[s];
        if (len < 0) return CS_E_INVALID_VALUE;
        if (len > (int64_t(1) << maxb)) return CS_E_CAPACITY;
        mx = std::max(mx, len);
        const int b = bucket_of(len);
        bkt[size_t(s)] = static_cast<unsigned char>(b);
        ++count[b];
    }
    if (max_len) *max_len = mx;
    int64_t start[kMaxBuckets] = {};
    for (int b = maxb, at = 0; b >= 0; --b) {
        start[b] = at;
        at += int(count[b]);
    }
    list.assign(size_t(segs), 0);
    {
        int64_t fill[kMaxBuckets];
        std::copy(start, start + kMaxBuckets, fill);
        for (int64_t s = 0; s < segs; ++s) list[size_t(fill[bkt[size_t(s)]]++)] = int(s);
    }
    items.clear();
    items.reserve(static_cast<size_t>(segs));
    for (int b = maxb; b >= 0; --b) {
        if (bucket_segments) bucket_segments[b] = count[b];
        int v = force_variant ? force_variant[b] : tuned_variant(dtype, b);
        if (v < 0 || !k.variant_ok(b, v)) v = default_variant(k, b);
        if (!k.variant_ok(b, v)) v = 2;
        if (variants) variants[b] = v;
        if (count[b] == 0) continue;
        const int64_t spi = k.segments_per_item(b, v);
        for (int64_t i = 0; i < count[b]; i += spi) {
            SegItem it;
            it.bucket = b;
            it.variant = v;
            it.first = int(start[b] + i);
            it.count = int(std::min(count[b] - i, spi));
            items.push_back(it);
        }
    }
    return CS_OK;
}

int seg_persistent_grid(int dtype, int which, int nitems) {
    static std::mutex mu;
    static int cache[8][2] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (dtype < 0 || dtype >= 8 || which < 0 || which > 1) return 0;
    if (!cache[dtype][which]) {
        int dev = 0, cus = 0, per_cu = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, seg_kernels_for(dtype).persistent_fn[which],
                                                              kSegBlock, 0);
        if (e != hipSuccess) { (void)hipGetLastError(); return 0; }
        cache[dtype][which] = std::max(1, cus * std::max(per_cu, 1));
    }
    return std::max(1, std::min(cache[dtype][which], std::max(nitems, 1)));
}

SegShape seg_shape(int dtype, const std::vector<SegItem>& items) {
    SegShape sh;
    const int sm = seg_kernels_for(dtype).small_max;
    sh.split = 0;
    while (sh.split < int(items.size()) && items[size_t(sh.split)].bucket > sm) ++sh.split;
    sh.grid[kSegMulti] = seg_persistent_grid(dtype, kSegMulti, sh.split);
    sh.grid[kSegSingle] = seg_persistent_grid(dtype, kSegSingle, int(items.size()) - sh.split);
    return sh;
}

hipError_t seg_launch_all(int dtype, const SegShape& sh, const void* in, void* out, const int64_t* off,
                          const int* d_list, const SegItem* d_items, int nitems, int* d_counters, hipStream_t s,
                          const int* d_range, bool reset) {
    if (!d_range && nitems <= 0) return hipSuccess;
    if (reset) hipLaunchKernelGGL(k_seg_reset, dim3(1), dim3(64), 0, s, d_counters);
    hipError_t e = hipGetLastError();
    const SegKernels k = seg_kernels_for(dtype);
    if (e == hipSuccess)
        e = k.persistent(kSegMulti, in, out, off, d_list, d_items, 0, sh.split, d_counters, d_range,
                         sh.grid[kSegMulti], s);
    if (e == hipSuccess)
        e = k.persistent(kSegSingle, in, out, off, d_list, d_items, sh.split, nitems, d_counters + 1, d_range,
                         sh.grid[kSegSingle], s);
    return e;
}

int ragged_prepare(int dtype, const int64_t* h_off, int64_t segs, RaggedPrep* p) {
    const LargeKernels lk = large_kernels_for(dtype);
    const int64_t chunk = lk.chunk;
    p->chunk_off.assign(1, h_off[0]);
    p->tiles.clear();
    p->passes = 0;
    for (int64_t s = 0; s < segs; ++s) {
        const int64_t b0 = h_off[s], len = h_off[s + 1] - b0;
        if (len < 0) return CS_E_INVALID_VALUE;
        if (len == 0) { p->chunk_off.push_back(b0); continue; }
        for (int64_t c = 0; c < len; c += chunk) p->chunk_off.push_back(b0 + std::min(len, c + chunk));
        if (len > chunk) {
            int q = 0;
            for (int64_t run = chunk; run < len; run *= 2) ++q;
            p->passes = std::max(p->passes, q);
            for (int64_t d0 = 0; d0 < len; d0 += lk.tile) p->tiles.push_back(MergeTile{b0, len, d0});
        }
    }
    const int rc = seg_build_items(dtype, p->chunk_off.data(), int64_t(p->chunk_off.size()) - 1, nullptr, p->list,
                                   p->items, nullptr, nullptr, nullptr);
    if (rc != CS_OK) return rc;
    p->shape = seg_shape(dtype, p->items);
    return p->shape.grid[0] > 0 && p->shape.grid[1] > 0 ? CS_OK : CS_E_NO_DEVICE;
}

hipError_t ragged_launch(int dtype, const void* in, void* out, void* scratch, const int64_t* d_chunk_off,
                         const int* d_list, const SegItem* d_items, int nitems, int* d_counters,
                         const SegShape& shape, MergeTile* d_tiles, int ntiles, int passes, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    hipError_t e = seg_launch_all(dtype, shape, in, out, d_chunk_off, d_list, d_items, nitems, d_counters, s);
    if (e != hipSuccess || ntiles <= 0 || passes <= 0) return e;
    const PartKernels pk = part_kernels_for(dtype);
    const int64_t chunk = large_kernels_for(dtype).chunk;
    void* a = out;
    void* b = scratch;
    for (int q = 0; e == hipSuccess && q < passes; ++q) {
        e = pk.seg_merge(a, b, d_tiles, ntiles, chunk << q, 0, s);
        std::swap(a, b);
    }
    if (e == hipSuccess && a != out) e = pk.seg_merge(a, out, d_tiles, ntiles, 0, 1, s);
    return e;
}

}  // namespace cs

extern "C" {

int cs_segmented_plan_create(cs_segplan** plan, int dtype, const int64_t* h_offsets, int64_t segments,
                             uint32_t flags) {
    return make_plan_impl(plan, dtype, h_offsets, segments, flags, nullptr);
}

int cs_segmented_plan_destroy(cs_segplan* plan) {
    free_plan(plan);
    return CS_OK;
}

int cs_segmented_sort(cs_segplan* plan, const void* in, void* out, const int64_t* d_offsets, hipStream_t stream) {
    return run_plan(plan, in, out, d_offsets, stream);
}

int cs_segmented_sort_pairs(cs_segplan* plan, const void* keys_in, const int32_t* vals_in, void* keys_out,
                            int32_t* vals_out, const int64_t* d_offsets, hipStream_t stream) {
    if (!plan) return CS_E_INVALID_VALUE;
    const int kd = plan->dtype == CS_REC_I32_I32 ? CS_I32 : (plan->dtype == CS_REC_I64_I32 ? CS_I64 : -1);
    if (kd < 0 || pair_record_dtype(kd) != plan->dtype) return CS_E_INVALID_VALUE;
    if (plan->total == 0) return CS_OK;
    if (!keys_in || !vals_in || !keys_out || !vals_out || !d_offsets) return CS_E_INVALID_VALUE;
    // the records sit at the segments' own positions: [base, base + total)
    const size_t esz = cs_dtype_size(plan->dtype);
    const size_t ksz = cs_dtype_size(kd);
    const int64_t end = plan->base + plan->total;
    char* rec = nullptr;
    hipError_t e = pool_malloc_async(reinterpret_cast<void**>(&rec), size_t(end) * esz, stream);
    if (e != hipSuccess) return int(e);
    const int64_t b = plan->base;
    e = pairs_pack(kd, static_cast<const char*>(keys_in) + b * ksz, vals_in + b, rec + b * esz, plan->total, stream);
    int rc = e == hipSuccess ? run_plan(plan, rec, rec, d_offsets, stream) : int(e);
    if (rc == CS_OK) {
        e = pairs_unpack(kd, rec + b * esz, static_cast<char*>(keys_out) + b * ksz, vals_out + b, plan->total, stream);
        if (e != hipSuccess) rc = int(e);
    }
    const hipError_t f = hipFreeAsync(rec, stream);
    return rc == CS_OK && f != hipSuccess ? int(f) : rc;
}

int cs_segmented_plan_info(const cs_segplan* plan, int32_t* work_items, int32_t* grid, int64_t* max_len,
                           int32_t* variants, int64_t* bucket_segments, int32_t buckets) {
    if (!plan) return CS_E_INVALID_VALUE;
    if (work_items) *work_items = plan->nitems;
    if (grid) *grid = plan->shape.grid[kSegMulti] * (plan->shape.split > 0) +
                      plan->shape.grid[kSegSingle] * (plan->nitems > plan->shape.split);
    if (max_len) *max_len = plan->max_len;
    for (int b = 0; b < buckets && b < kMaxBuckets; ++b) {
        if (variants) variants[b] = plan->variant[b];
        if (bucket_segments) bucket_segments[b] = plan->bucket_segments[b];
    }
    return CS_OK;
}

int32_t cs_segmented_max_len(int dtype) {
    if (dtype < CS_I32 || dtype > CS_F64) return 0;
    return int32_t(1) << seg_kernels_for(dtype).max_logw;
}

int cs_segmented_sort_plain(int dtype, const void* in, void* out, const int64_t* d_offsets, int64_t segments,
                            int32_t max_len, hipStream_t stream) {
    if (dtype < CS_I32 || dtype > CS_F64 || segments < 0 || max_len < 0) return CS_E_INVALID_VALUE;
    if (segments == 0) return CS_OK;
    if (!in || !out || !d_offsets) return CS_E_INVALID_VALUE;
    const SegKernels k = seg_kernels_for(dtype);
    if (max_len > (int32_t(1) << k.max_logw)) return CS_E_CAPACITY;
    const hipError_t e = k.plain(in, out, d_offsets, segments, bucket_of(max_len), stream);
    return e == hipSuccess ? CS_OK : int(e);
}

int cs_segmented_get_variant(int dtype, int32_t bucket, int32_t* variant) {
    if (dtype < CS_I32 || dtype > CS_F64 || bucket < 0 || bucket >= kMaxBuckets || !variant) return CS_E_INVALID_VALUE;
    *variant = tuned_variant(dtype, bucket);
    return CS_OK;
}

int cs_segmented_set_variant(int dtype, int32_t bucket, int32_t variant) {
    if (dtype < CS_I32 || dtype > CS_F64 || bucket < 0 || bucket >= kMaxBuckets) return CS_E_INVALID_VALUE;
    if (variant != -1 && !seg_kernels_for(dtype).variant_ok(bucket, variant)) return CS_E_UNSUPPORTED;
    set_tuned_variant(dtype, bucket, variant);
    return CS_OK;
}

// Tuner: for every bucket, time each variant on a batch of full-length
// segments (2^b elements each, ~2^22 elements in total, uniform keys from
// the library's generator) through the persistent kernel, keep the fastest
// (median of `trials` timed runs).  Allocates scratch; not for timed paths.
int cs_segmented_tune(int dtype, int32_t trials, hipStream_t stream) {
    if (dtype < CS_I32 || dtype > CS_F64 || trials < 1) return CS_E_INVALID_VALUE;
    const SegKernels k = seg_kernels_for(dtype);
    const size_t esz = cs_dtype_size(dtype);
    const int64_t budget = int64_t(1) << 22;
    void *in = nullptr, *out = nullptr;
    int64_t* d_off = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipError_t e = hipMalloc(&in, size_t(budget) * esz);
    if (e == hipSuccess) e = hipMalloc(&out, size_t(budget) * esz);
    if (e == hipSuccess) e = hipMalloc(&d_off, size_t(budget + 1) * sizeof(int64_t));
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e == hipSuccess) e = generate(dtype, CS_DIST_UNIFORM, budget, budget, 12345, in, stream);
    int rc = e == hipSuccess ? CS_OK : int(e);
    for (int b = 1; b <= k.max_logw && rc == CS_OK; ++b) {
        const int64_t len = int64_t(1) << b;
        const int64_t segs = budget / len;
        std::vector<int64_t> off(size_t(segs) + 1);
        for (int64_t s = 0; s <= segs; ++s) off[size_t(s)] = s * len;
        e = hipMemcpy(d_off, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) { rc = int(e); break; }
        float best = 0;
        int best_v = -1;
        for (int v = 0; v < kSegVariants; ++v) {
            if (!k.variant_ok(b, v)) continue;
            int force[kMaxBuckets];
            for (int& f : force) f = v;
            cs_segplan* p = nullptr;
            rc = make_plan_impl(&p, dtype, off.data(), segs, CS_SEG_NO_GRAPH, force);
            if (rc != CS_OK) break;
            std::vector<float> ms;
            for (int r = 0; r <= trials && rc == CS_OK; ++r) {
                (void)hipEventRecord(t0, stream);
                rc = run_plan(p, in, out, d_off, stream);
                (void)hipEventRecord(t1, stream);
                (void)hipEventSynchronize(t1);
                float m = 0;
                (void)hipEventElapsedTime(&m, t0, t1);
                if (r > 0) ms.push_back(m);   // first run warms up
            }
            free_plan(p);
            if (rc != CS_OK) break;
            std::sort(ms.begin(), ms.end());
            const float med = ms[ms.size() / 2];
            if (best_v < 0 || med < best) { best = med; best_v = v; }
        }
        if (rc == CS_OK && best_v >= 0) set_tuned_variant(dtype, b, best_v);
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    if (d_off) (void)hipFree(d_off);
    return rc;
}

}  // extern "C"
