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
//  5. scatter: the workgroup re-classifies its block tile by tile, stages
//     each tile in bucket order in LDS (ranks from LDS atomics; order inside
//     a bucket is free, the sort is unstable like the reference's) and
//     writes each bucket's run of the tile to consecutive addresses.
// Two levels run steps 2-5 twice, the second once per first-level bucket,
// with that bucket's splitters drawn from its own sample.  The final
// buckets are then sorted as one ragged batch planned on the device (the
// reference reads its counters back to the host instead, cusort.hpp:
// 1967-1977): buckets just over the single-wave tile as head and tail
// merged afterwards, buckets longer than a chunk by chunk sorts and merge
// passes.  Nothing is read back to the host.
#pragma once

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "large.hpp"
#include "segmented.hpp"

namespace cs {

constexpr int kPartThreads = 512;        // threads of the count kernel
constexpr int kPartMaxSplitters = 2047;          // 2P-1 <= 4095 buckets (12-bit bucket ids)

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


// Elements per thread per tile of the count and scatter kernels.
template <int DT> constexpr int part_items() { return sizeof(typename Elem<DT>::Storage) == 4 ? 16 : 8; }
// The scatter stages larger tiles (16384 4-byte / 8192 8-byte / 4096 16-byte
// elements) with a wider workgroup: each bucket's run per tile is then long
// enough that its writes fill whole cache lines.
template <int DT> constexpr int part_scatter_threads() {
    return sizeof(typename Elem<DT>::Storage) >= 16 ? 512 : 1024;
}
template <int DT> constexpr int part_scatter_tile() { return part_items<DT>() * part_scatter_threads<DT>(); }

// Work of one workgroup of the count and scatter kernels: the element range
// [begin, end) of one segment, classified against the splitters
// sp[sp0, sp0 + ns).  The segment's count matrix is bucket-major with one
// column per workgroup of the segment: (bucket b, this workgroup) is cell
// `cell + b * stride`.  lo / hi index the splitters that bound every key of
// the segment from below / above (-1: unbounded); they only narrow the
// lookup table.  A one-level partition is a single segment over the whole
// array; the second level runs one segment per first-level bucket.
struct PartBlock {
    int64_t begin;
    int64_t end;
    int64_t cell;
    int32_t stride;
    int32_t sp0;
    int32_t ns;
    int32_t lo;
    int32_t hi;
    int32_t pad;
};

// Bucket start positions of a segment: for bucket b of the segment,
// starts[first + b] = scanned count of cell (cell0 + b * stride).
struct PartSeg {
    int64_t cell0;
    int64_t first;
    int32_t stride;
    int32_t nb;
    int32_t equal;   // every bucket of the segment holds one key value (a first-level equality bucket)
    int32_t pad;
};

// Splitters and the first level's work list, from the sorted sample: the
// P - 1 global splitters are every (m / P)-th sample; with two levels the
// p1 - 1 first-level splitters are every p2-th global splitter (global
// splitter (j + 1) p2 - 1), so the second level of first-level bucket j uses
// the p2 - 1 global splitters inside it.  The first level is one segment
// over the whole array cut into workgroups of block_elems elements.
template <int DT>
__global__ void k_part_prep(const typename Elem<DT>::Storage* sorted, int m, int parts, int p1, int p2,
                            typename Elem<DT>::Key* gsp, typename Elem<DT>::Key* l1sp, int64_t n,
                            int64_t block_elems, int64_t blocks1, PartBlock* work1, PartSeg* segs1) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < parts - 1) gsp[j] = Elem<DT>::key(sorted[(j + 1) * m / parts - 1]);
    if (p2 > 1 && j < p1 - 1) l1sp[j] = Elem<DT>::key(sorted[(j + 1) * p2 * m / parts - 1]);
    const int ns = p1 - 1;
    if (j == 0) segs1[0] = PartSeg{0, 0, int32_t(blocks1), 2 * ns + 1, 0, 0};
    if (j < blocks1) {
        const int64_t x0 = j * block_elems;
        work1[j] = PartBlock{x0, x0 + block_elems < n ? x0 + block_elems : n, j, int32_t(blocks1), 0, ns, -1, -1, 0};
    }
}

// Second-level splitters drawn from each first-level bucket (two-stage
// sampling): first-level part j (bucket 2j, keys strictly between
// first-level splitters j-1 and j) of len keys gets p2 = part_p2(len)
// second-level parts; s2 * p2 keys sampled at hashed positions of its own
// range in the first level's output are sorted as one chunk (the rest of
// the chunk padded with the maximum key), and every s2-th of them becomes
// one of its p2 - 1 splitters.  The splitter list keeps the layout of one
// list with a stride of p2cap entries per part: part j's second-level
// splitters at j * p2cap .., first-level splitter j at (j + 1) * p2cap - 1
// (the entries between repeat it), so the second level is laid out and run
// as with one global list.  Each bucket's sample follows its actual size and
// key range, so its final buckets come out at the target size whatever size
// the first level gave it.
// starts: the first level's scanned count matrix (bucket b starts at
// starts[b * blocks1]).
// Second-level parts of a first-level part of len keys (two-stage
// sampling): as many as keep its buckets at `target` or above (so a caller's
// part bound, folded into target, still holds), at least 1, at most p2cap.
__host__ __device__ __forceinline__ int part_p2(int64_t len, int64_t target, int p2cap) {
    const int64_t p = len / target;
    return int(p < 1 ? 1 : (p > p2cap ? p2cap : p));
}

template <int DT>
__device__ __forceinline__ void part_l2_range(const int64_t* starts, int64_t blocks1, int p1, int64_t n, int j,
                                              int64_t& b0, int64_t& len) {
    const int b = 2 * j;
    b0 = starts[int64_t(b) * blocks1];
    const int64_t e = b + 1 < 2 * p1 - 1 ? starts[int64_t(b + 1) * blocks1] : n;
    len = e - b0;
}

// smp[j * chunk + i]: i < s2 * p2 sampled keys of part j (none if the part
// is empty), the rest of the chunk the maximum key.  One thread per slot.
template <int DT>
__global__ void k_part_l2_gather(const typename Elem<DT>::Storage* src, const int64_t* starts, int64_t blocks1,
                                 int p1, int64_t target, int p2cap, int s2, int chunk, int64_t n,
                                 typename Elem<DT>::Storage* smp) {
    using E = Elem<DT>;
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= int64_t(p1) * chunk) return;
    const int j = int(q / chunk), i = int(q - int64_t(j) * chunk);
    int64_t b0 = 0, len = 0;
    part_l2_range<DT>(starts, blocks1, p1, n, j, b0, len);
    if (len > 0 && i < s2 * part_p2(len, target, p2cap)) {
        const uint32_t x = uint32_t(i) * 2654435761u ^ uint32_t(j) * 0x9e3779b9u;
        const uint64_t h = (uint64_t(part_hash(x)) << 32) | part_hash(x ^ 0x85ebca6bu);
        smp[q] = src[b0 + int64_t(h % uint64_t(len))];
    } else {
        smp[q] = E::make(E::kPad, INT32_MAX);
    }
}

// gsp[j * p2cap + k]: part j's second-level splitters (k < p2 - 1) from its
// sorted sample, then first-level splitter j repeated up to (j + 1) * p2cap
// - 1 (the last part: the maximum key), so the list never decreases.  One
// thread per entry.
template <int DT>
__global__ void k_part_l2_pick(const typename Elem<DT>::Storage* smp, const int64_t* starts, int64_t blocks1, int p1,
                               int64_t target, int p2cap, int s2, int chunk, int64_t n,
                               const typename Elem<DT>::Key* l1sp, typename Elem<DT>::Key* gsp) {
    using E = Elem<DT>;
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= int64_t(p1) * p2cap - 1) return;
    const int j = int(q / p2cap), k = int(q - int64_t(j) * p2cap);
    int64_t b0 = 0, len = 0;
    part_l2_range<DT>(starts, blocks1, p1, n, j, b0, len);
    const int p2 = len > 0 ? part_p2(len, target, p2cap) : 1;
    if (k < p2 - 1) gsp[q] = E::key(smp[int64_t(j) * chunk + int64_t(k + 1) * s2 - 1]);
    else gsp[q] = j < p1 - 1 ? l1sp[j] : E::kPad;
}

// Classification.  A lookup table indexed by kPartLutBits bits of the key's
// order-preserving unsigned image, taken relative to the segment's lower
// bound and scaled to its width, narrows the splitter search to the
// splitters that share the key's prefix (any prefix that does not decrease
// as the key grows keeps this exact).  With 4096 entries and a few hundred
// splitters per segment, spread-out keys meet zero or one candidate; for
// keys crowded under one prefix the range is the full list and the search
// is the plain binary search.  Built by every workgroup from the splitters
// in LDS.
constexpr int kPartLutBits = 12;
constexpr int kPartLutSize = (1 << kPartLutBits) + 1;

template <class K> __device__ __forceinline__ uint64_t part_image(K k) {
    if constexpr (sizeof(K) == 4) {
        return std::is_signed<K>::value ? uint64_t(uint32_t(k) ^ 0x80000000u) : uint64_t(uint32_t(k));
    } else {
        return std::is_signed<K>::value ? uint64_t(k) ^ (uint64_t(1) << 63) : uint64_t(k);
    }
}

struct PartLutMap {   // prefix = min((image - base) >> shift, 2^kPartLutBits - 1), 0 below base
    uint64_t base;
    int shift;
};

template <class K> __device__ __forceinline__ uint32_t part_prefix(K k, PartLutMap m) {
    if constexpr (sizeof(K) == 4) {   // 32-bit arithmetic: base and shift fit
        const uint32_t u = uint32_t(part_image(k)), b = uint32_t(m.base);
        const uint32_t r = __builtin_elementwise_sub_sat(u, b) >> m.shift;   // 0 below the base
        return r < uint32_t(kPartLutSize - 1) ? r : uint32_t(kPartLutSize - 2);
    } else {
        const uint64_t u = part_image(k);
        const uint64_t r = u > m.base ? (u - m.base) >> m.shift : 0;
        return r < uint64_t(kPartLutSize - 1) ? uint32_t(r) : uint32_t(kPartLutSize - 2);
    }
}

// Map of a segment whose keys lie in [image(sp[lo]), image(sp[hi])] (whole
// key range where unbounded).
template <class K> __device__ __forceinline__ PartLutMap part_lut_map(const K* sp_all, int lo, int hi) {
    const uint64_t top = sizeof(K) == 4 ? 0xFFFFFFFFull : ~uint64_t(0);
    const uint64_t base = lo >= 0 ? part_image(sp_all[lo]) : 0;
    const uint64_t end = hi >= 0 ? part_image(sp_all[hi]) : top;
    const uint64_t range = end > base ? end - base : 0;
    const int bits = range ? 64 - __clzll(range) : 0;
    return PartLutMap{base, bits > kPartLutBits ? bits - kPartLutBits : 0};
}

// Lookup-table entry of prefix x: range = lo | hi << 16, where lo splitters
// have a prefix below x and hi have a prefix at most x, so the splitters that
// can tie with or bracket a key of prefix x are sp[lo .. hi).  For 4-byte
// keys the entry also carries sp[lo]: a key with at most one candidate
// splitter (almost every key once the table is finer than the splitters) is
// then classified by one 8-byte LDS read and no other.
template <class K, bool kCarry = (sizeof(K) == 4)> struct alignas(8) PartLutEntry {
    uint32_t range;
    K first;
};
template <class K> struct PartLutEntry<K, false> {
    uint32_t range;
};

// Builds the table from the splitters sp[0, ns) (sorted, two readable
// entries past ns) with the workgroup's NT threads: a histogram of the
// splitters' prefixes, then one exclusive scan of it (lo(x) = sum of the
// counts below x, hi(x) = lo(x) + count(x)).  wsum holds NT / 64 ints.
template <int NT, class K>
__device__ __forceinline__ void part_build_lut(const K* sp, int ns, PartLutEntry<K>* lut, PartLutMap m, int* wsum) {
    constexpr int X = kPartLutSize - 1;          // entries
    constexpr int C = X / NT;                    // entries per thread, contiguous
    static_assert(X % NT == 0 && NT % 64 == 0, "table entries split evenly over the workgroup");
    const int tid = int(threadIdx.x);
    const int lane = tid & 63, wave = tid >> 6;
    for (int x = tid; x < X; x += NT) lut[x].range = 0;
    __syncthreads();
    for (int i = tid; i < ns; i += NT) atomicAdd(&lut[part_prefix(sp[i], m)].range, 1u);
    __syncthreads();
    int c[C];
    int sum = 0;
    #pragma unroll
    for (int q = 0; q < C; ++q) {
        c[q] = int(lut[tid * C + q].range);
        sum += c[q];
    }
    int incl = sum;
    #pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int lo = incl - sum;
    for (int q = 0; q < wave; ++q) lo += wsum[q];
    #pragma unroll
    for (int q = 0; q < C; ++q) {
        PartLutEntry<K> e;
        e.range = uint32_t(lo) | (uint32_t(lo + c[q]) << 16);
        if constexpr (sizeof(K) == 4) e.first = sp[lo];   // lo <= ns: readable
        lut[tid * C + q] = e;
        lo += c[q];
    }
}

// Bucket of key k: j = number of splitters <= k; a key equal to splitter
// j-1 goes to the equality bucket 2j-1, any other key to bucket 2j.  The
// splitters below the key's candidate range are below it and those above
// are above it (the prefix does not decrease as the key grows), and a key
// left of every candidate cannot equal splitter lo-1, whose prefix is
// smaller than its own.  One candidate is compared in registers (4-byte
// keys: taken from the table entry); a longer candidate range (keys crowded
// under one prefix) is binary-searched.
template <class E>
__device__ __forceinline__ int part_bucket(typename E::Key k, const typename E::Key* sp,
                                           const PartLutEntry<typename E::Key>* lut, PartLutMap m) {
    using K = typename E::Key;
    const uint32_t x = part_prefix(k, m);
    uint32_t range;
    K s{};
    if constexpr (sizeof(K) == 4) {   // range and first candidate in one 8-byte read
        const uint2 raw = reinterpret_cast<const uint2*>(lut)[x];
        range = raw.x;
        __builtin_memcpy(&s, &raw.y, 4);
    } else {
        range = lut[x].range;
    }
    const int lo = int(range & 0xFFFFu), hi = int(range >> 16);
    if constexpr (sizeof(K) != 4) {
        if (hi - lo == 1) s = sp[lo];
    }
    // zero or one candidate, without a branch
    const bool ge = hi > lo && !(k < s);
    int j = lo + int(ge);
    bool eq = ge && s == k;
    if (hi - lo > 1) {   // several candidates: binary search
        j = lo;
        int b = hi;
        while (j < b) {
            const int mid = (j + b) >> 1;
            if (!(k < sp[mid])) j = mid + 1;
            else b = mid;
        }
        eq = j > lo && sp[j - 1] == k;
    }
    return 2 * j - int(eq);
}

// Adds `one` (1 per valid lane) to counter[b] for every valid lane and
// returns the lane's previous value plus its rank among the wave's lanes of
// the same bucket.  When every valid lane of the wave has the same bucket
// (sorted or heavily repeated input) one atomic serves the whole wave:
// per-lane atomics on one address would serialise.
__device__ __forceinline__ int part_wave_add(int* counter, int b, bool valid) {
    const int lane = int(__lane_id());
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return 0;
    // lane reads with a wave-uniform index: v_readlane, no LDS traffic
    const int first = __ffsll(static_cast<unsigned long long>(vmask)) - 1;
    const int b0 = __builtin_amdgcn_readlane(b, first);
    const uint64_t smask = __ballot(valid && b == b0);
    if (smask == vmask) {   // wave-uniform branch
        int base = 0;
        if (lane == first) base = atomicAdd(&counter[b0], __popcll(vmask));
        base = __builtin_amdgcn_readlane(base, first);
        return base + __popcll(vmask & ((uint64_t(1) << lane) - 1));
    }
    return valid ? atomicAdd(&counter[b], 1) : 0;
}

// Raw buffer loads of a workgroup's element range: a position at or past
// `elems` reads as zero instead of leaving the range, so a tile's loads need
// no clamping, and their addresses are 32-bit offsets from a scalar base
// (a range is at most 512 KiB).
template <class S>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t part_rsrc(const S* base, int elems) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<S*>(base), short(0), (elems > 0 ? elems : 0) * int(sizeof(S)),
                                             0x00020000);   // raw (untyped, unswizzled) buffer
}

template <class S> __device__ __forceinline__ S part_load(__amdgpu_buffer_rsrc_t r, int i) {
    const int off = i * int(sizeof(S));
    S s;
    if constexpr (sizeof(S) == 4) {
        const int v = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
        __builtin_memcpy(&s, &v, 4);
    } else if constexpr (sizeof(S) == 8) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        __builtin_memcpy(&s, &v, 8);
    } else {
        static_assert(sizeof(S) == 16, "4-, 8- or 16-byte elements");
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        __builtin_memcpy(&s, &v, 16);
    }
    return s;
}

// Per-workgroup histogram: every thread classifies part_items keys per tile,
// their loads all in flight together (coalesced across the workgroup).
// Dynamic LDS of k_part_count for up to ns splitters per workgroup:
// splitters (+2 readable entries) and the histogram.
template <int DT> size_t part_count_lds(int ns) {
    return ((size_t(ns + 2) * sizeof(typename Elem<DT>::Key) + 15) & ~size_t(15)) + size_t(2 * ns + 1) * 4;
}

template <int DT>
__global__ void __launch_bounds__(kPartThreads)
k_part_count(const typename Elem<DT>::Storage* src, const typename Elem<DT>::Key* splitters, const PartBlock* work,
             int max_ns, int64_t* counts, const int64_t* nwork) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    if (nwork && int64_t(blockIdx.x) >= *nwork) return;   // grid sized by a bound, work list built on the device
    constexpr int IT = part_items<DT>();
    constexpr int NT = kPartThreads;
    extern __shared__ __align__(16) unsigned char part_lds[];
    typename E::Key* sp = reinterpret_cast<typename E::Key*>(part_lds);
    int* hist = reinterpret_cast<int*>(part_lds + ((size_t(max_ns + 2) * sizeof(typename E::Key) + 15) & ~size_t(15)));
    __shared__ PartLutEntry<typename E::Key> lut[kPartLutSize - 1];
    __shared__ int wsum[NT / 64];
    __shared__ int crowded;   // some bucket drew a quarter of the first tile
    const PartBlock w = work[blockIdx.x];
    const int ns = w.ns;
    const int nb = 2 * ns + 1;
    const int tid = int(threadIdx.x);
    if (tid == 0) crowded = 0;
    for (int i = tid; i < ns + 2; i += NT) sp[i] = i < ns ? splitters[w.sp0 + i] : typename E::Key{};
    for (int i = tid; i < nb; i += NT) hist[i] = 0;
    const PartLutMap map = part_lut_map(splitters, w.lo, w.hi);
    __syncthreads();
    part_build_lut<NT>(sp, ns, lut, map, wsum);
    __syncthreads();
    // 32-bit element indices relative to the workgroup's range (at most
    // 512 KiB long).  The next tile's loads are issued before this
    // tile is classified; loads past the end read zero (their results are
    // masked below).
    const S* base = src + w.begin;
    const int len = int(w.end - w.begin);
    constexpr int STEP = IT * NT;
    bool agg = false;
    S cur[IT];
    {
        const __amdgpu_buffer_rsrc_t r = part_rsrc(base, len);
        #pragma unroll
        for (int k = 0; k < IT; ++k) cur[k] = part_load<S>(r, k * NT + tid);
    }
    for (int t0 = 0; t0 < len; t0 += STEP) {
        S nxt[IT];
        const __amdgpu_buffer_rsrc_t r = part_rsrc(base + t0 + STEP, len - t0 - STEP);
        #pragma unroll
        for (int k = 0; k < IT; ++k) nxt[k] = part_load<S>(r, k * NT + tid);
        if (agg) {   // workgroup-uniform: sorted or repeated keys crowd one counter
            #pragma unroll
            for (int k = 0; k < IT; ++k) {
                const bool valid = t0 + k * NT + tid < len;
                const int b = valid ? part_bucket<E>(E::key(cur[k]), sp, lut, map) : 0;
                (void)part_wave_add(hist, b, valid);
            }
        } else if (t0 + STEP <= len) {   // a whole tile: no lane is masked
            #pragma unroll
            for (int k = 0; k < IT; ++k) atomicAdd(&hist[part_bucket<E>(E::key(cur[k]), sp, lut, map)], 1);
        } else {
            #pragma unroll
            for (int k = 0; k < IT; ++k)
                if (t0 + k * NT + tid < len) atomicAdd(&hist[part_bucket<E>(E::key(cur[k]), sp, lut, map)], 1);
        }
        #pragma unroll
        for (int k = 0; k < IT; ++k) cur[k] = nxt[k];
        if (t0 == 0) {   // after the first tile: switch to wave-aggregated counting if crowded
            __syncthreads();
            for (int b = tid; b < nb; b += NT)
                if (hist[b] >= STEP / 4) crowded = 1;
            __syncthreads();
            agg = crowded != 0;
        }
    }
    __syncthreads();
    for (int b = tid; b < nb; b += NT) counts[w.cell + int64_t(b) * w.stride] = hist[b];
}

// Dynamic LDS of k_part_scatter for up to ns splitters per workgroup:
// splitters, tile histogram / offsets, the workgroup's running output
// positions, the tile staged in bucket order and the bucket of every staged
// element.
constexpr size_t part_align(size_t x) { return (x + 15) & ~size_t(15); }
template <int DT> constexpr size_t part_scatter_lds(int ns) {
    using E = Elem<DT>;
    const size_t nb = size_t(2 * ns + 1);
    return part_align(size_t(ns + 2) * sizeof(typename E::Key)) + part_align(nb * 4) + 2 * part_align(nb * 8) +
           part_align(size_t(part_scatter_tile<DT>()) * sizeof(typename E::Storage)) +
           part_align(size_t(part_scatter_tile<DT>()) * 2) + 64;
}

// Most splitters per workgroup for which the scatter fits a CU's 160 KB of
// LDS: its dynamic carve-up, the lookup table and a margin for the small
// static arrays.  Both levels' part counts are capped by it (past it, the
// buckets grow and the rare one longer than a tile takes the long path).
template <int DT> constexpr int part_max_splitters() {
    constexpr size_t kLds = 160 * 1024;
    constexpr size_t fixed = sizeof(PartLutEntry<typename Elem<DT>::Key>) * (kPartLutSize - 1) + 256;
    int ns = kPartMaxSplitters;
    while (ns > 1 && part_scatter_lds<DT>(ns) + fixed > kLds) --ns;
    return ns;
}

// Scatter.  Per tile of part_scatter_tile elements: classify, rank inside
// the tile with LDS atomics, scan the tile histogram (which also advances
// the workgroup's running output position of every bucket), stage the tile
// in bucket order in LDS, then write it out with consecutive threads on
// consecutive positions of each bucket's run (coalesced where the
// reference's per-element atomics scatter every key to its own address).
// max_ns sizes the LDS carve-up (the largest ns of any workgroup of the
// launch).
template <int DT>
__global__ void __launch_bounds__(part_scatter_threads<DT>())
k_part_scatter(const typename Elem<DT>::Storage* src, const typename Elem<DT>::Key* splitters, const PartBlock* work,
               int max_ns, const int64_t* offsets, typename Elem<DT>::Storage* dst, const int64_t* nwork) {
    using E = Elem<DT>;
    if (nwork && int64_t(blockIdx.x) >= *nwork) return;   // grid sized by a bound, work list built on the device
    using K = typename E::Key;
    using S = typename E::Storage;
    constexpr int IT = part_items<DT>();
    constexpr int NT = part_scatter_threads<DT>();
    constexpr int TILE = part_scatter_tile<DT>();
    extern __shared__ __align__(16) unsigned char part_lds[];
    const PartBlock w = work[blockIdx.x];
    const int ns = w.ns;
    const int nb = 2 * ns + 1;
    const int max_nb = 2 * max_ns + 1;
    // LDS: splitters (+2 readable entries), tile histogram / offsets, running
    // output position and this tile's write base (run - offset) per bucket,
    // the staged tile and the bucket of every staged element
    K* sp = reinterpret_cast<K*>(part_lds);
    int* toff = reinterpret_cast<int*>(part_lds + part_align(size_t(max_ns + 2) * sizeof(K)));
    int64_t* run = reinterpret_cast<int64_t*>(reinterpret_cast<unsigned char*>(toff) + part_align(size_t(max_nb) * 4));
    int64_t* wbase = run + part_align(size_t(max_nb) * 8) / 8;
    S* stage = reinterpret_cast<S*>(reinterpret_cast<unsigned char*>(wbase) + part_align(size_t(max_nb) * 8));
    uint16_t* sbkt = reinterpret_cast<uint16_t*>(reinterpret_cast<unsigned char*>(stage) +
                                                 part_align(size_t(TILE) * sizeof(S)));
    __shared__ int wsum[NT / 64];
    __shared__ PartLutEntry<K> lut[kPartLutSize - 1];
    __shared__ int crowded;   // a bucket drew a quarter of some tile: rank with wave-aggregated atomics
    const int tid = int(threadIdx.x);
    const int lane = tid & 63, wave = tid >> 6;
    if (tid == 0) crowded = 0;
    for (int i = tid; i < ns + 2; i += NT) sp[i] = i < ns ? splitters[w.sp0 + i] : K{};
    for (int b = tid; b < nb; b += NT) {
        toff[b] = 0;
        run[b] = offsets[w.cell + int64_t(b) * w.stride];
    }
    const PartLutMap map = part_lut_map(splitters, w.lo, w.hi);
    __syncthreads();
    part_build_lut<NT>(sp, ns, lut, map, wsum);
    __syncthreads();
    const int per_thread = (nb + NT - 1) / NT;   // scan chunk of each thread
    // 32-bit indices relative to the workgroup's range; the next tile's loads
    // are issued before this tile is processed (loads past the end read zero,
    // as in the count kernel)
    const S* base = src + w.begin;
    const int len = int(w.end - w.begin);
    S elem[IT];
    {
        const __amdgpu_buffer_rsrc_t r = part_rsrc(base, len);
        #pragma unroll
        for (int k = 0; k < IT; ++k) elem[k] = part_load<S>(r, k * NT + tid);
    }
    for (int t0 = 0; t0 < len; t0 += TILE) {
        const int tn = len - t0 < TILE ? len - t0 : TILE;
        S nxt[IT];
        const __amdgpu_buffer_rsrc_t r = part_rsrc(base + t0 + TILE, len - t0 - TILE);
        #pragma unroll
        for (int k = 0; k < IT; ++k) nxt[k] = part_load<S>(r, k * NT + tid);
        uint32_t br[IT];   // bucket (12 bits) | rank inside the tile << 12
        if (crowded) {   // workgroup-uniform (set before the last barrier)
            #pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int x = k * NT + tid;
                const int b = x < tn ? part_bucket<E>(E::key(elem[k]), sp, lut, map) : 0;
                const int r = part_wave_add(toff, b, x < tn);
                br[k] = uint32_t(b) | (uint32_t(r) << 12);
            }
        } else if (tn == TILE) {   // a whole tile: no lane is masked
            #pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int b = part_bucket<E>(E::key(elem[k]), sp, lut, map);
                br[k] = uint32_t(b) | (uint32_t(atomicAdd(&toff[b], 1)) << 12);
            }
        } else {
            #pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int x = k * NT + tid;
                const int b = x < tn ? part_bucket<E>(E::key(elem[k]), sp, lut, map) : 0;
                const int r = x < tn ? atomicAdd(&toff[b], 1) : 0;
                br[k] = uint32_t(b) | (uint32_t(r) << 12);
            }
        }
        __syncthreads();
        // exclusive scan of the tile histogram in place: a chunk per thread,
        // then a wave scan of the chunk sums and the wave totals
        const int c0 = tid * per_thread;
        int csum = 0;
        for (int b = c0; b < c0 + per_thread && b < nb; ++b) csum += toff[b];
        int incl = csum;
        #pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int acc = incl - csum;
        for (int q = 0; q < wave; ++q) acc += wsum[q];
        for (int b = c0; b < c0 + per_thread && b < nb; ++b) {
            const int c = toff[b];
            if (c >= TILE / 4) crowded = 1;   // read by the next tile, after three barriers
            toff[b] = acc;
            wbase[b] = run[b] - acc;   // staged position j of bucket b goes to wbase[b] + j
            run[b] += c;
            acc += c;
        }
        __syncthreads();
        #pragma unroll
        for (int k = 0; k < IT; ++k) {
            if (k * NT + tid < tn) {
                const int b = int(br[k] & 0xFFFu);
                const int pos = toff[b] + int(br[k] >> 12);
                stage[pos] = elem[k];
                sbkt[pos] = uint16_t(b);
            }
        }
        __syncthreads();
        for (int j = tid; j < tn; j += NT) dst[wbase[sbkt[j]] + j] = stage[j];
        for (int b = tid; b < nb; b += NT) toff[b] = 0;   // offsets were last read before the barrier above
        #pragma unroll
        for (int k = 0; k < IT; ++k) elem[k] = nxt[k];
        __syncthreads();
    }
}

// Bucket starts of every segment (one workgroup per segment), and whether
// each bucket is an equality bucket (every key in it equal: already sorted,
// tie order being free).
// Workgroup 0 also clears zero[0, nzero) (the final plan's counters).
__global__ void k_part_starts(const PartSeg* segs, const int64_t* offsets, int64_t* starts, unsigned char* equal,
                              int* zero, int nzero);


// Exclusive scan of int64 values in place, three launches: per-block sums,
// a single-block scan of those sums, and the per-block scan with its offset.
// With d_n the count is *d_n (at most n, known on the device).
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void k_scan_sums(const int64_t* v, int64_t n, int64_t* sums, const int64_t* d_n);
__global__ void k_scan_top(int64_t* sums, int64_t nblocks);
__global__ void k_scan_apply(int64_t* v, int64_t n, const int64_t* sums, const int64_t* d_n);

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


// Buckets longer than the on-chip tile (keys crowded into a range the sample
// missed; equality buckets never get here, the final sort runs in place and
// leaves them as they are) are sorted without the host knowing whether any
// exist, by two ordinary launches that return at once when there are none:
//  * k_part_long_chunks: the chunk sorts of every long bucket, chunk q of the
//    running count over the buckets going to workgroup q mod grid (src ->
//    out, the chunk kernel's network);
//  * k_part_long_merge: one workgroup per long bucket runs all of its merge
//    passes (out <-> other) and the copy back after an odd number of them:
//    per pass it finds the co-ranks of all the bucket's merge tiles (one
//    search per thread, all in flight together), then merges the tiles in
//    turn.  Only workgroup barriers: no workgroup ever waits for another.
// A sample that misses a crowded key range leaves a few buckets just over
// the tile; only an input built against the hashed sample positions would
// make one long bucket large, and that bucket is then merged by one CU.
// long_list[0, *nlong): the long buckets, any order; st: bucket starts.
template <int DT>
__global__ void __launch_bounds__(kMergeThreads)
k_part_long_chunks(const typename Elem<DT>::Storage* src, typename Elem<DT>::Storage* out, const int64_t* st,
                   const int* long_list, const int* nlong_p) {
    using E = Elem<DT>;
    constexpr int LW = chunk_logw<DT>();
    constexpr int LE = auto_loge(DT, LW);
    using C = TileCfg<E, LW, LE>;
    static_assert(C::BLOCK == kMergeThreads && C::APB == 1, "one chunk per workgroup of the merge width");
    constexpr int64_t CHUNK = int64_t(1) << LW;
    __shared__ typename E::Key lds_k[C::LDS_KEYS];
    __shared__ int32_t lds_p[C::LDS_PAYLOAD];
    const int nlong = *nlong_p;
    const int64_t G = gridDim.x, me = blockIdx.x;
    int64_t base = 0;   // chunks of the buckets before l (the same in every workgroup)
    for (int l = 0; l < nlong; ++l) {
        const int b = long_list[l];
        const int64_t b0 = st[b], len = st[b + 1] - b0;
        const int64_t nch = (len + CHUNK - 1) / CHUNK;
        for (int64_t c = ((me - base) % G + G) % G; c < nch; c += G) {
            const int64_t c0 = c * CHUNK;
            const int clen = int(len - c0 < CHUNK ? len - c0 : CHUNK);
            tile_sort<E, LW, LE>(src + b0 + c0, out + b0 + c0, clen, int(threadIdx.x), lds_k, lds_p);
            __syncthreads();   // the tile's LDS is reused by the next chunk
        }
        base += nch;
    }
}

// cor: scratch of one int64 per merge tile of all long buckets together
// (tile_base: the tiles of the buckets before l, the same running count in
// every workgroup).
template <int DT>
__global__ void __launch_bounds__(kMergeThreads)
k_part_long_merge(typename Elem<DT>::Storage* out, typename Elem<DT>::Storage* other, const int64_t* st,
                  const int* long_list, const int* nlong_p, int64_t* cor) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    constexpr int64_t CHUNK = int64_t(1) << chunk_logw<DT>();
    constexpr int TILE = merge_tile<DT>();
    const int nlong = *nlong_p;
    const int tid = int(threadIdx.x);
    int64_t tile_base = 0;
    for (int l = 0; l < nlong; ++l) {
        const int b = long_list[l];
        const int64_t b0 = st[b], len = st[b + 1] - b0;
        const int64_t ntiles = (len + TILE - 1) / TILE;
        if (l % int(gridDim.x) == int(blockIdx.x)) {
            int64_t* tc = cor + tile_base;
            S* a = out + b0;
            S* b2 = other + b0;
            for (int64_t run = CHUNK; run < len; run *= 2) {
                for (int64_t q = tid; q < ntiles; q += kMergeThreads) {
                    const MergeGeom g = merge_geom<TILE>(len, run, q * TILE);
                    const S* A = a + g.base;
                    tc[q] = thread_corank<E>(A, g.na, A + g.na, g.nb, g.d0);
                }
                __syncthreads();
                for (int64_t q = 0; q < ntiles; ++q) {
                    const MergeGeom g = merge_geom<TILE>(len, run, q * TILE);
                    // a tile that does not end its run pair finds its end
                    // co-rank as the next tile's start
                    merge_tile_body<DT>(a, b2, len, run, q * TILE, tc[q], merge_end_corank(g, tc + q + 1));
                    __syncthreads();   // LDS reused by the next tile
                }
                S* t = a;
                a = b2;
                b2 = t;
            }
            if (a != out + b0)   // odd number of passes: back into out
                for (int64_t i = tid; i < len; i += kMergeThreads) out[b0 + i] = a[i];
            __syncthreads();
        }
        tile_base += ntiles;
    }
}

// Buckets a little longer than the single-wave tile (2^SM < len <= 2^(SM+1),
// SM = seg_small_max) are sorted by the persistent kernel as two segments,
// head [0, 2^SM) and tail [2^SM, len), both in single-wave tiles, instead of
// one padded tile of 2^(SM+1) spread over several waves; this kernel then
// merges the two sorted halves of each such bucket in place.  The whole
// bucket is staged in LDS before anything is written back, every thread
// finds its co-rank there and merges its outputs into registers (left half
// first on equal keys), and the result leaves through LDS with consecutive
// threads on consecutive addresses.  A persistent grid walks the list, each
// workgroup loading its next bucket while it merges the current one.
template <int DT>
__global__ void __launch_bounds__(kMergeThreads)
k_part_split_merge(typename Elem<DT>::Storage* out, const int64_t* st, const int* split_list, const int* nsplit_p) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    constexpr int H = 1 << seg_small_max<DT>();
    constexpr int W = 2 * H;
    constexpr int IPT = W / kMergeThreads;
    static_assert(W % kMergeThreads == 0, "whole outputs per thread");
    constexpr int PIN = sizeof(S) == 4 ? 3 : 2;   // input pads, as in the merge passes
    constexpr int PAD = sizeof(S) == 4 ? 5 : (sizeof(S) == 8 ? 4 : 3);
    constexpr int LDS_IN = W + (W >> PIN);
    constexpr int LDS_OUT = W + (W >> PAD);
    __shared__ S lds[LDS_IN > LDS_OUT ? LDS_IN : LDS_OUT];
    const int nsplit = *nsplit_p;
    const int G = int(gridDim.x);
    const int tid = int(threadIdx.x);
    // the next bucket's elements are loaded into registers while this one
    // is merged
    S nxt[IPT];
    int64_t nb0 = 0;
    int ntotal = 0;
    auto fetch = [&](int q) {
        const int b = split_list[q];
        nb0 = st[b];
        ntotal = int(st[b + 1] - nb0);   // H < ntotal <= W
        #pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int x = tid + k * kMergeThreads;
            if (x < ntotal) nxt[k] = out[nb0 + x];
        }
    };
    if (int(blockIdx.x) < nsplit) fetch(int(blockIdx.x));
    for (int q = int(blockIdx.x); q < nsplit; q += G) {
        const int total = ntotal;
        S* base = out + nb0;
        #pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int x = tid + k * kMergeThreads;
            if (x < total) lds[x + (x >> PIN)] = nxt[k];
        }
        __syncthreads();
        if (q + G < nsplit) fetch(q + G);
        const int la = H, lb = total - H;
        const int dt0 = tid * IPT < total ? tid * IPT : total;
        int i = lds_corank<E, PIN>(lds, la, lb, dt0);
        int j = dt0 - i;
        const int ia0 = i < total ? i : 0;
        const int jb0 = la + j < total ? la + j : 0;
        S cur_a = lds[ia0 + (ia0 >> PIN)];
        S cur_b = lds[jb0 + (jb0 >> PIN)];
        S outv[IPT];
        #pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const bool take_a = j >= lb || (i < la && !(E::key(cur_b) < E::key(cur_a)));
            outv[k] = take_a ? cur_a : cur_b;
            i += take_a ? 1 : 0;
            j += take_a ? 0 : 1;
            int r = take_a ? i : la + j;   // refill of the side taken (unused once it is exhausted)
            r = r < total ? r : 0;
            const S nv = lds[r + (r >> PIN)];
            cur_a = take_a ? nv : cur_a;
            cur_b = take_a ? cur_b : nv;
        }
        __syncthreads();
        #pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const int x = dt0 + k;
            if (x < total) lds[x + (x >> PAD)] = outv[k];
        }
        __syncthreads();
        for (int x = tid; x < total; x += kMergeThreads) base[x] = lds[x + (x >> PAD)];
        __syncthreads();   // LDS reused by the next bucket
    }
}

struct PartKernels {
    hipError_t (*sample)(const void* src, int64_t n, void* samples, int m, hipStream_t s) = nullptr;
    // k_part_prep
    hipError_t (*prep)(const void* sorted_samples, int m, int parts, int p1, int p2, void* gsp, void* l1sp, int64_t n,
                       int64_t block_elems, int64_t blocks1, PartBlock* work1, PartSeg* segs1,
                       hipStream_t s) = nullptr;
    // k_part_l2_gather + k_part_l2_pick around the caller's chunk sort
    hipError_t (*l2_gather)(const void* src, const int64_t* starts, int64_t blocks1, int p1, int64_t target,
                            int p2cap, int s2, int chunk, int64_t n, void* smp, hipStream_t s) = nullptr;
    hipError_t (*l2_pick)(const void* smp, const int64_t* starts, int64_t blocks1, int p1, int64_t target, int p2cap,
                          int s2, int chunk, int64_t n, const void* l1sp, void* gsp, hipStream_t s) = nullptr;
    // count / scatter over `blocks` work items; with nwork (device) only the
    // first *nwork of them exist and the rest of the grid returns at once.
    hipError_t (*count)(const void* src, const void* splitters, const PartBlock* work, int blocks, int max_ns,
                        int64_t* counts, const int64_t* nwork, hipStream_t s) = nullptr;
    hipError_t (*scatter)(const void* src, const void* splitters, const PartBlock* work, int blocks, int max_ns,
                          const int64_t* offsets, void* dst, const int64_t* nwork, hipStream_t s) = nullptr;
    // k_part_long_chunks + k_part_long_merge over the long buckets of a
    // device plan (grids of at most grid_cap workgroups).
    hipError_t (*long_sort)(const void* src, void* out, void* other, const int64_t* st, const int* long_list,
                            const int* nlong, int64_t* cor, int grid_cap, hipStream_t s) = nullptr;
    // k_part_split_merge over the split buckets of a device plan.
    hipError_t (*split_merge)(void* out, const int64_t* st, const int* split_list, const int* nsplit, int grid_cap,
                              hipStream_t s) = nullptr;
    // Co-rank search + merge pass (or a plain copy with copy_only); writes
    // the co-ranks into items.
    hipError_t (*seg_merge)(const void* src, void* dst, MergeTile* items, int nitems, int64_t run,
                            int copy_only, hipStream_t s) = nullptr;
    int key_size = 0;
    int max_splitters = 0;   // part_max_splitters: splitters per workgroup the scatter's LDS holds
};

template <int DT>
PartKernels part_kernels() {
    using E = Elem<DT>;
    using S = typename E::Storage;
    using K = typename E::Key;
    PartKernels k;
    k.key_size = int(sizeof(K));
    k.max_splitters = part_max_splitters<DT>();
    k.sample = [](const void* src, int64_t n, void* samples, int m, hipStream_t s) {
        hipLaunchKernelGGL(k_part_sample<DT>, dim3(unsigned((m + 255) / 256)), dim3(256), 0, s,
                           static_cast<const S*>(src), n, static_cast<S*>(samples), m);
        return hipGetLastError();
    };
    k.prep = [](const void* sorted, int m, int parts, int p1, int p2, void* gsp, void* l1sp, int64_t n,
                int64_t block_elems, int64_t blocks1, PartBlock* work1, PartSeg* segs1, hipStream_t s) {
        const int64_t threads = std::max<int64_t>(parts, blocks1);
        hipLaunchKernelGGL(k_part_prep<DT>, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s,
                           static_cast<const S*>(sorted), m, parts, p1, p2, static_cast<K*>(gsp),
                           static_cast<K*>(l1sp), n, block_elems, blocks1, work1, segs1);
        return hipGetLastError();
    };
    k.l2_gather = [](const void* src, const int64_t* starts, int64_t blocks1, int p1, int64_t target, int p2cap,
                     int s2, int chunk, int64_t n, void* smp, hipStream_t s) {
        const int64_t threads = int64_t(p1) * chunk;
        hipLaunchKernelGGL(k_part_l2_gather<DT>, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s,
                           static_cast<const S*>(src), starts, blocks1, p1, target, p2cap, s2, chunk, n,
                           static_cast<S*>(smp));
        return hipGetLastError();
    };
    k.l2_pick = [](const void* smp, const int64_t* starts, int64_t blocks1, int p1, int64_t target, int p2cap, int s2,
                   int chunk, int64_t n, const void* l1sp, void* gsp, hipStream_t s) {
        const int64_t threads = int64_t(p1) * p2cap;
        hipLaunchKernelGGL(k_part_l2_pick<DT>, dim3(unsigned((threads + 255) / 256)), dim3(256), 0, s,
                           static_cast<const S*>(smp), starts, blocks1, p1, target, p2cap, s2, chunk, n,
                           static_cast<const K*>(l1sp), static_cast<K*>(gsp));
        return hipGetLastError();
    };
    k.count = [](const void* src, const void* sp, const PartBlock* work, int blocks, int max_ns, int64_t* counts,
                 const int64_t* nwork, hipStream_t s) {
        hipLaunchKernelGGL(k_part_count<DT>, dim3(unsigned(blocks)), dim3(kPartThreads), part_count_lds<DT>(max_ns),
                           s, static_cast<const S*>(src), static_cast<const K*>(sp), work, max_ns, counts, nwork);
        return hipGetLastError();
    };
    k.scatter = [](const void* src, const void* sp, const PartBlock* work, int blocks, int max_ns,
                   const int64_t* offsets, void* dst, const int64_t* nwork, hipStream_t s) {
        const size_t lds = part_scatter_lds<DT>(max_ns);
        if (lds > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_part_scatter<DT>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_part_scatter<DT>, dim3(unsigned(blocks)), dim3(part_scatter_threads<DT>()), lds, s,
                           static_cast<const S*>(src), static_cast<const K*>(sp), work, max_ns, offsets,
                           static_cast<S*>(dst), nwork);
        return hipGetLastError();
    };
    k.split_merge = [](void* out, const int64_t* st, const int* split_list, const int* nsplit, int grid_cap,
                       hipStream_t s) {
        // eight workgroups per CU walk the list (any grid is correct)
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        int grid = std::max(1, 8 * cus);
        if (grid_cap > 0 && grid > grid_cap) grid = grid_cap;
        hipLaunchKernelGGL(k_part_split_merge<DT>, dim3(unsigned(grid)), dim3(kMergeThreads), 0, s,
                           static_cast<S*>(out), st, split_list, nsplit);
        return hipGetLastError();
    };
    k.long_sort = [](const void* src, void* out, void* other, const int64_t* st, const int* long_list,
                     const int* nlong, int64_t* cor, int grid_cap, hipStream_t s) {
        // one workgroup per CU for each kernel (any grid is correct)
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        int grid = std::max(1, cus);
        if (grid_cap > 0 && grid > grid_cap) grid = grid_cap;
        hipLaunchKernelGGL(k_part_long_chunks<DT>, dim3(unsigned(grid)), dim3(kMergeThreads), 0, s,
                           static_cast<const S*>(src), static_cast<S*>(out), st, long_list, nlong);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_part_long_merge<DT>, dim3(unsigned(grid)), dim3(kMergeThreads), 0, s,
                           static_cast<S*>(out), static_cast<S*>(other), st, long_list, nlong, cor);
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
