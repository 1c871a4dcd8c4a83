// Sample-sort partition driver (see partition.hpp) and the prefix-scan
// kernels it uses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/cusort_hip.h"
#include "partition.hpp"

namespace cs {

__global__ void __launch_bounds__(kScanThreads) k_scan_sums(const int64_t* v, int64_t n, int64_t* sums) {
    __shared__ int64_t part[kScanThreads];
    const int64_t b0 = int64_t(blockIdx.x) * kScanTile;
    int64_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = b0 + int64_t(k) * kScanThreads + threadIdx.x;
        if (i < n) s += v[i];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = kScanThreads / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

// Exclusive scan of `nblocks` sums by one workgroup (sequential chunks).
__global__ void __launch_bounds__(kScanThreads) k_scan_top(int64_t* sums, int64_t nblocks) {
    __shared__ int64_t buf[kScanThreads];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < nblocks; c0 += kScanThreads) {
        const int64_t i = c0 + threadIdx.x;
        const int64_t v = i < nblocks ? sums[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (int w = 1; w < kScanThreads; w <<= 1) {   // Hillis-Steele inclusive scan
            const int64_t add = int(threadIdx.x) >= w ? buf[threadIdx.x - w] : 0;
            __syncthreads();
            buf[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nblocks) sums[i] = carry + buf[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kScanThreads - 1) carry += buf[kScanThreads - 1];
        __syncthreads();
    }
}

// Exclusive scan of each tile, offset by the tile's scanned sum.
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(int64_t* v, int64_t n, const int64_t* sums) {
    __shared__ int64_t buf[kScanThreads];
    const int64_t b0 = int64_t(blockIdx.x) * kScanTile;
    int64_t vals[kScanItems];
    int64_t s = 0;
    for (int k = 0; k < kScanItems; ++k) {           // thread owns kScanItems consecutive values
        const int64_t i = b0 + int64_t(threadIdx.x) * kScanItems + k;
        vals[k] = i < n ? v[i] : 0;
        s += vals[k];
    }
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int w = 1; w < kScanThreads; w <<= 1) {
        const int64_t add = int(threadIdx.x) >= w ? buf[threadIdx.x - w] : 0;
        __syncthreads();
        buf[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t run = sums[blockIdx.x] + buf[threadIdx.x] - s;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = b0 + int64_t(threadIdx.x) * kScanItems + k;
        if (i < n) v[i] = run;
        run += vals[k];
    }
}

__global__ void k_bucket_totals(const int64_t* offsets, int nb, int blocks, int64_t n, int64_t* bucket_off) {
    const int b = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (b < nb) bucket_off[b] = offsets[int64_t(b) * blocks];
    if (b == nb) bucket_off[b] = n;
}

PartKernels part_kernels_i32();
PartKernels part_kernels_u32();
PartKernels part_kernels_f32();
PartKernels part_kernels_u64();
PartKernels part_kernels_i64();
PartKernels part_kernels_f64();
PartKernels part_kernels_rec32();
PartKernels part_kernels_rec64();

PartKernels part_kernels_for(int dtype) {
    switch (dtype) {
        case CS_I32: return part_kernels_i32();
        case CS_U32: return part_kernels_u32();
        case CS_F32: return part_kernels_f32();
        case CS_U64: return part_kernels_u64();
        case CS_I64: return part_kernels_i64();
        case CS_F64: return part_kernels_f64();
        case CS_REC_I32_I32: return part_kernels_rec32();
        case CS_REC_I64_I32: return part_kernels_rec64();
        default: return PartKernels{};
    }
}

namespace {

hipError_t exclusive_scan(int64_t* v, int64_t n, int64_t* sums, hipStream_t s) {
    const int64_t blocks = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_sums, dim3(unsigned(blocks)), dim3(kScanThreads), 0, s, v, n, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, sums, blocks);
    hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(blocks)), dim3(kScanThreads), 0, s, v, n, sums);
    return hipGetLastError();
}

template <class T>
struct Scoped {   // stream-ordered temporary
    T* p = nullptr;
    hipStream_t s;
    explicit Scoped(hipStream_t st) : s(st) {}
    hipError_t alloc(size_t count) { return hipMallocAsync(reinterpret_cast<void**>(&p), count * sizeof(T), s); }
    ~Scoped() { if (p) (void)hipFreeAsync(p, s); }
};

}  // namespace

// Sorts data[0, n) into out (out may equal data) with tmp as scratch of at
// least n elements.  Two host synchronisations: bucket sizes after the
// scatter, and before the one-shot segment plan is released.
int partition_sort(int dtype, void* data, int64_t n, void* out, void* tmp, size_t tmp_elems, int min_parts,
                   int max_parts, hipStream_t stream) {
    const LargeKernels lk = large_kernels_for(dtype);
    const PartKernels pk = part_kernels_for(dtype);
    const size_t esz = cs_dtype_size(dtype);
    if (tmp_elems < size_t(n)) return CS_E_WORKSPACE;
    if (n <= lk.chunk) return cs_batched_sort(dtype, data, out, 1, int32_t(n), CS_METHOD_AUTO, 0, 0, stream);

    // Default: buckets of ~2^18 keys when the chunk is 2^13 keys (4-byte
    // elements), ~2^17 when it is 2^12.  The largest bucket then lands just
    // above a power of two and needs an even number of merge passes (6), so
    // the result ends where it started and no copy-back pass is needed; more
    // parts fragment the scatter's runs.  At most 1024 parts.  Measured on
    // MI355X, 2^28 int32: 8.0 ms at 1024 parts, 8.8 at 512 (7 passes + a
    // copy), 10.6 at 2048; 2^27 uint64: 9.1 ms at 1024, 9.7 at 256.
    const int bucket_log2 = lk.chunk == 8192 ? 18 : 17;
    int parts = int(std::min<int64_t>(1024, std::max<int64_t>(2, n >> bucket_log2)));
    if (min_parts > 0) parts = std::max(parts, min_parts);   // the caller's MinNbParts / MaxNbParts
    if (max_parts > 0) parts = std::min(parts, max_parts);
    parts = std::max(2, std::min(parts, kPartMaxSplitters + 1));
    const int ns = parts - 1;
    const int nb = 2 * ns + 1;
    const int m = parts * 64;   // oversampling: 64 sample keys per part
    const int blocks = int(std::min<int64_t>(1024, std::max<int64_t>(1, (n + 65535) / 65536)));
    const int64_t per_block = (n + blocks - 1) / blocks;
    const int64_t cells = int64_t(nb) * blocks;

    Scoped<unsigned char> samples(stream), samples_tmp(stream), splitters(stream);
    Scoped<int64_t> counts(stream), sums(stream), bucket_off(stream);
    hipError_t e = samples.alloc(size_t(m) * esz);
    if (e == hipSuccess) e = samples_tmp.alloc(size_t(m) * esz);
    if (e == hipSuccess) e = splitters.alloc(size_t(std::max(ns, 1)) * size_t(pk.key_size));
    if (e == hipSuccess) e = counts.alloc(size_t(cells));
    if (e == hipSuccess) e = sums.alloc(size_t((cells + kScanTile - 1) / kScanTile) + 1);
    if (e == hipSuccess) e = bucket_off.alloc(size_t(nb) + 1);
    if (e != hipSuccess) return int(e);

    // 1-2. splitters from a sorted hashed sample
    e = pk.sample(data, n, samples.p, m, stream);
    if (e != hipSuccess) return int(e);
    int rc = cs_sort_merge(dtype, samples.p, m, 0, 0, samples_tmp.p, size_t(m), 1, stream);
    if (rc != CS_OK) return rc;
    e = pk.splitters(samples.p, m, parts, splitters.p, stream);
    // 3-5. count, scan, scatter (data -> tmp)
    if (e == hipSuccess) e = pk.count(data, n, splitters.p, ns, per_block, blocks, counts.p, stream);
    if (e == hipSuccess) e = exclusive_scan(counts.p, cells, sums.p, stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bucket_totals, dim3(unsigned((nb + 1 + 255) / 256)), dim3(256), 0, stream, counts.p, nb,
                           blocks, n, bucket_off.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = pk.scatter(data, n, splitters.p, ns, per_block, blocks, counts.p, tmp, stream);
    std::vector<int64_t> boff(size_t(nb) + 1);
    if (e == hipSuccess)
        e = hipMemcpyAsync(boff.data(), bucket_off.p, boff.size() * sizeof(int64_t), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return int(e);

    // 6. one ragged batch over all buckets; long buckets cut into chunks
    const int64_t chunk = lk.chunk;
    std::vector<int64_t> seg{0};
    std::vector<MergeTile> tiles;
    int passes = 0;
    for (int b = 0; b < nb; ++b) {
        const int64_t b0 = boff[size_t(b)], len = boff[size_t(b) + 1] - b0;
        if (len <= 0) continue;
        for (int64_t c = 0; c < len; c += chunk) seg.push_back(b0 + std::min(len, c + chunk));
        if (len > chunk && (b % 2) == 0) {   // equality buckets (odd) are sorted once their chunks are
            int p = 0;
            for (int64_t run = chunk; run < len; run *= 2) ++p;
            passes = std::max(passes, p);
            for (int64_t d0 = 0; d0 < len; d0 += lk.tile) tiles.push_back(MergeTile{b0, len, d0});
        }
    }
    cs_segplan* plan = nullptr;
    rc = cs_segmented_plan_create(&plan, dtype, seg.data(), int64_t(seg.size()) - 1, CS_SEG_NO_GRAPH);
    if (rc != CS_OK) return rc;
    Scoped<int64_t> d_seg(stream);
    Scoped<MergeTile> d_tiles(stream);
    e = d_seg.alloc(seg.size());
    if (e == hipSuccess) e = hipMemcpyAsync(d_seg.p, seg.data(), seg.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && !tiles.empty()) e = d_tiles.alloc(tiles.size());
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpyAsync(d_tiles.p, tiles.data(), tiles.size() * sizeof(MergeTile), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) rc = cs_segmented_sort(plan, tmp, out, d_seg.p, stream);
    else rc = int(e);
    // 7. merge passes over the long buckets (out <-> tmp), then final placement
    void* a = out;
    void* b2 = tmp;
    for (int p = 0; rc == CS_OK && p < passes; ++p) {
        e = pk.seg_merge(a, b2, d_tiles.p, int(tiles.size()), chunk << p, 0, stream);
        if (e != hipSuccess) rc = int(e);
        std::swap(a, b2);
    }
    if (rc == CS_OK && a != out) {
        e = pk.seg_merge(a, out, d_tiles.p, int(tiles.size()), 0, 1, stream);
        if (e != hipSuccess) rc = int(e);
    }
    const hipError_t se = hipStreamSynchronize(stream);   // the plan's buffers are in use until here
    cs_segmented_plan_destroy(plan);
    if (rc == CS_OK && se != hipSuccess) rc = int(se);
    return rc;
}

}  // namespace cs
