// Sample-sort partitioning for the large-array Sort / HybridSort paths
// (reference: Sort_core_buildPivots cusort.hpp:1745-1800, CountParts
// 1322-1375, BuildPrefix 1700-1740, MoveElements 1377-1417; the V3 variants
// 1420-1693 with mono-valued bucket handling 1497-1502, 2323-2330).
//
//  1. sample: M keys at hashed positions, sorted on the device; S = P-1
//     splitters are every (M/P)-th sample.
//  2. classify: j = #splitters <= key; a key equal to splitter j-1 goes to
//     the equality bucket 2j-1, any other key to bucket 2j (2P-1 buckets).
//     Repeated splitters (heavy duplicates) therefore collect every copy of
//     that key in one equality bucket, which is already sorted.
//  3. count: one workgroup per contiguous block of the input builds an LDS
//     histogram and writes column (block) of a bucket-major count matrix.
//  4. scan: exclusive prefix sum of the matrix (bucket-major, block-minor)
//     gives each (bucket, block) its output offset.
//  5. scatter: the workgroup re-classifies its block and places every
//     element at its bucket's next slot (LDS atomics; order inside a bucket
//     is free, the sort is unstable like the reference's).
// The host then reads the bucket sizes once (the reference also reads its
// counters back, cusort.hpp:1967-1977) and sorts the buckets as one ragged
// batch; buckets longer than the on-chip tile are cut into chunks whose
// sorted runs are merged by tiled merge passes over a tile list.
#pragma once

#include "large.hpp"

namespace cs {

constexpr int kPartThreads = 256;
constexpr int kPartMaxSplitters = 2047;          // 2P-1 <= 4095 buckets
constexpr int kPartMaxBuckets = 2 * (kPartMaxSplitters + 1) - 1;

__device__ __forceinline__ uint32_t part_hash(uint32_t x) {   // PCG output permutation of x
    x = x * 747796405u + 2891336453u;
    const uint32_t w = ((x >> ((x >> 28u) + 4u)) ^ x) * 277803737u;
    return (w >> 22u) ^ w;
}

template <int DT>
__global__ void k_part_sample(const typename Elem<DT>::Storage* src, int64_t n, typename Elem<DT>::Storage* samples,
                              int m) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < m) {
        const uint64_t h = (uint64_t(part_hash(uint32_t(i))) << 32) | part_hash(uint32_t(i) ^ 0x9e3779b9u);
        samples[i] = src[int64_t(h % uint64_t(n))];
    }
}

template <int DT>
__global__ void k_part_splitters(const typename Elem<DT>::Storage* sorted_samples, int m, int parts,
                                 typename Elem<DT>::Key* splitters) {
    const int j = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (j < parts - 1) splitters[j] = Elem<DT>::key(sorted_samples[int64_t(j + 1) * m / parts - 1]);
}

template <class E>
__device__ __forceinline__ int part_bucket(typename E::Key k, const typename E::Key* sp, int ns) {
    int lo = 0, hi = ns;                 // j = number of splitters <= k
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!(k < sp[mid])) lo = mid + 1;
        else hi = mid;
    }
    return (lo > 0 && !(sp[lo - 1] < k)) ? 2 * lo - 1 : 2 * lo;
}

template <int DT>
__global__ void __launch_bounds__(kPartThreads)
k_part_count(const typename Elem<DT>::Storage* src, int64_t n, const typename Elem<DT>::Key* splitters, int ns,
             int64_t per_block, int blocks, int64_t* counts) {
    using E = Elem<DT>;
    __shared__ typename E::Key sp[kPartMaxSplitters];
    __shared__ int hist[kPartMaxBuckets];
    const int nb = 2 * ns + 1;
    for (int i = int(threadIdx.x); i < ns; i += kPartThreads) sp[i] = splitters[i];
    for (int i = int(threadIdx.x); i < nb; i += kPartThreads) hist[i] = 0;
    __syncthreads();
    const int64_t b0 = int64_t(blockIdx.x) * per_block;
    const int64_t b1 = b0 + per_block < n ? b0 + per_block : n;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kPartThreads)
        atomicAdd(&hist[part_bucket<E>(E::key(src[i]), sp, ns)], 1);
    __syncthreads();
    for (int b = int(threadIdx.x); b < nb; b += kPartThreads) counts[int64_t(b) * blocks + blockIdx.x] = hist[b];
}

template <int DT>
__global__ void __launch_bounds__(kPartThreads)
k_part_scatter(const typename Elem<DT>::Storage* src, int64_t n, const typename Elem<DT>::Key* splitters, int ns,
               int64_t per_block, int blocks, const int64_t* offsets, typename Elem<DT>::Storage* dst) {
    using E = Elem<DT>;
    __shared__ typename E::Key sp[kPartMaxSplitters];
    __shared__ unsigned long long next[kPartMaxBuckets];
    const int nb = 2 * ns + 1;
    for (int i = int(threadIdx.x); i < ns; i += kPartThreads) sp[i] = splitters[i];
    for (int b = int(threadIdx.x); b < nb; b += kPartThreads)
        next[b] = (unsigned long long)offsets[int64_t(b) * blocks + blockIdx.x];
    __syncthreads();
    const int64_t b0 = int64_t(blockIdx.x) * per_block;
    const int64_t b1 = b0 + per_block < n ? b0 + per_block : n;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += kPartThreads) {
        const typename E::Storage s = src[i];
        const int b = part_bucket<E>(E::key(s), sp, ns);
        const unsigned long long pos = atomicAdd(&next[b], 1ull);
        dst[pos] = s;
    }
}

// Exclusive scan of int64 values in place, three launches: per-block sums,
// a single-block scan of those sums, and the per-block scan with its offset.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void k_scan_sums(const int64_t* v, int64_t n, int64_t* sums);
__global__ void k_scan_top(int64_t* sums, int64_t nblocks);
__global__ void k_scan_apply(int64_t* v, int64_t n, const int64_t* sums);
__global__ void k_bucket_totals(const int64_t* offsets, int nb, int blocks, int64_t n, int64_t* bucket_off);

// Merge pass over a list of oversized buckets: item = (bucket base, bucket
// length, first output position of this tile inside the bucket), plus the
// co-rank of the tile's first output, which k_seg_merge_corank fills in
// before every pass.  The tiles of a bucket are consecutive items, so a
// tile that does not end its run pair finds its end co-rank in the next item.
// With copy_only the tile is copied unchanged (equality buckets, final
// placement after an odd number of passes).
struct MergeTile {
    int64_t base;
    int64_t len;
    int64_t d0;
    int64_t a0 = 0;   // co-rank of the tile's first output
};

template <int DT>
__global__ void __launch_bounds__(kCorankThreads)
k_seg_merge_corank(const typename Elem<DT>::Storage* src, MergeTile* items, int nitems, int64_t run) {
    using E = Elem<DT>;
    const int q = int(blockIdx.x) * kCorankThreads + int(threadIdx.x);
    if (q >= nitems) return;
    const MergeTile it = items[q];
    const MergeGeom g = merge_geom<merge_tile<DT>()>(it.len, run, it.d0);
    const typename E::Storage* A = src + it.base + g.base;
    items[q].a0 = thread_corank<E>(A, g.na, A + g.na, g.nb, g.d0);
}

template <int DT>
__global__ void __launch_bounds__(kMergeThreads)
k_seg_merge_pass(const typename Elem<DT>::Storage* src, typename Elem<DT>::Storage* dst, const MergeTile* items,
                 int64_t run, int copy_only) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    constexpr int TILE = merge_tile<DT>();
    const MergeTile it = items[blockIdx.x];
    const int tid = int(threadIdx.x);
    if (copy_only) {
        const int64_t end = it.d0 + TILE < it.len ? it.d0 + TILE : it.len;
        for (int64_t i = it.d0 + tid; i < end; i += kMergeThreads) dst[it.base + i] = src[it.base + i];
        return;
    }
    const MergeGeom g = merge_geom<TILE>(it.len, run, it.d0);
    merge_tile_body<DT>(src + it.base, dst + it.base, it.len, run, it.d0, it.a0,
                        merge_end_corank(g, &items[blockIdx.x + 1].a0));
}

struct PartKernels {
    hipError_t (*sample)(const void* src, int64_t n, void* samples, int m, hipStream_t s) = nullptr;
    hipError_t (*splitters)(const void* sorted_samples, int m, int parts, void* splitters, hipStream_t s) = nullptr;
    hipError_t (*count)(const void* src, int64_t n, const void* splitters, int ns, int64_t per_block, int blocks,
                        int64_t* counts, hipStream_t s) = nullptr;
    hipError_t (*scatter)(const void* src, int64_t n, const void* splitters, int ns, int64_t per_block, int blocks,
                          const int64_t* offsets, void* dst, hipStream_t s) = nullptr;
    // Co-rank search + merge pass (or a plain copy with copy_only); writes
    // the co-ranks into items.
    hipError_t (*seg_merge)(const void* src, void* dst, MergeTile* items, int nitems, int64_t run,
                            int copy_only, hipStream_t s) = nullptr;
    int key_size = 0;
};

template <int DT>
PartKernels part_kernels() {
    using E = Elem<DT>;
    using S = typename E::Storage;
    using K = typename E::Key;
    PartKernels k;
    k.key_size = int(sizeof(K));
    k.sample = [](const void* src, int64_t n, void* samples, int m, hipStream_t s) {
        hipLaunchKernelGGL(k_part_sample<DT>, dim3(unsigned((m + 255) / 256)), dim3(256), 0, s,
                           static_cast<const S*>(src), n, static_cast<S*>(samples), m);
        return hipGetLastError();
    };
    k.splitters = [](const void* sorted, int m, int parts, void* sp, hipStream_t s) {
        hipLaunchKernelGGL(k_part_splitters<DT>, dim3(unsigned((parts + 255) / 256)), dim3(256), 0, s,
                           static_cast<const S*>(sorted), m, parts, static_cast<K*>(sp));
        return hipGetLastError();
    };
    k.count = [](const void* src, int64_t n, const void* sp, int ns, int64_t per_block, int blocks, int64_t* counts,
                 hipStream_t s) {
        hipLaunchKernelGGL(k_part_count<DT>, dim3(unsigned(blocks)), dim3(kPartThreads), 0, s,
                           static_cast<const S*>(src), n, static_cast<const K*>(sp), ns, per_block, blocks, counts);
        return hipGetLastError();
    };
    k.scatter = [](const void* src, int64_t n, const void* sp, int ns, int64_t per_block, int blocks,
                   const int64_t* offsets, void* dst, hipStream_t s) {
        hipLaunchKernelGGL(k_part_scatter<DT>, dim3(unsigned(blocks)), dim3(kPartThreads), 0, s,
                           static_cast<const S*>(src), n, static_cast<const K*>(sp), ns, per_block, blocks, offsets,
                           static_cast<S*>(dst));
        return hipGetLastError();
    };
    k.seg_merge = [](const void* src, void* dst, MergeTile* items, int nitems, int64_t run, int copy_only,
                     hipStream_t s) {
        if (nitems <= 0) return hipSuccess;
        if (!copy_only) {
            const int64_t cblocks = (int64_t(nitems) + kCorankThreads - 1) / kCorankThreads;
            hipLaunchKernelGGL(k_seg_merge_corank<DT>, dim3(unsigned(cblocks)), dim3(kCorankThreads), 0, s,
                               static_cast<const S*>(src), items, nitems, run);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_seg_merge_pass<DT>, dim3(unsigned(nitems)), dim3(kMergeThreads), 0, s,
                           static_cast<const S*>(src), static_cast<S*>(dst), items, run, copy_only);
        return hipGetLastError();
    };
    return k;
}

PartKernels part_kernels_for(int dtype);

}  // namespace cs
