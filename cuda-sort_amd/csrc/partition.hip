// Sample-sort partition driver (see partition.hpp) and the prefix-scan
// kernels it uses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/cusort_hip.h"
#include "gen.hpp"
#include "ragged.hpp"
#include "scratch.hpp"

namespace cs {

__global__ void __launch_bounds__(kScanThreads) k_scan_sums(const int64_t* v, int64_t n, int64_t* sums,
                                                            const int64_t* d_n) {
    __shared__ int64_t part[kScanThreads];
    const int64_t b0 = int64_t(blockIdx.x) * kScanTile;
    if (d_n) {   // count known on the device; blocks past it sum to zero
        n = *d_n < n ? *d_n : n;
        if (b0 >= n) {
            if (threadIdx.x == 0) sums[blockIdx.x] = 0;
            return;
        }
    }
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
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(int64_t* v, int64_t n, const int64_t* sums,
                                                             const int64_t* d_n) {
    __shared__ int64_t buf[kScanThreads];
    const int64_t b0 = int64_t(blockIdx.x) * kScanTile;
    if (d_n) {
        n = *d_n < n ? *d_n : n;
        if (b0 >= n) return;
    }
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

__global__ void k_part_starts(const PartSeg* segs, const int64_t* offsets, int64_t* starts, unsigned char* equal,
                              int* zero, int nzero) {
    if (blockIdx.x == 0)
        for (int i = int(threadIdx.x); i < nzero; i += int(blockDim.x)) zero[i] = 0;
    const PartSeg g = segs[blockIdx.x];
    for (int b = int(threadIdx.x); b < g.nb; b += int(blockDim.x)) {
        starts[g.first + b] = offsets[g.cell0 + int64_t(b) * g.stride];
        equal[g.first + b] = (g.equal || (b & 1)) ? 1 : 0;
    }
}

// Second partition level, laid out on the device from the first level's
// scanned count matrix (bucket b starts at cell (b, 0) = counts1[b *
// blocks1]; no host round trip): first-level bucket b becomes segment b.
// An equality bucket (odd b) is carried through as one bucket.  With one
// global splitter list (p2cap = 0) a bucket no longer than `target` is
// carried through too and any other is split by the p2 - 1 splitters
// inside it (first-level splitter j is list entry (j + 1) p2 - 1); with
// two-stage sampling (p2cap > 0) part j is split by its own part_p2(len) - 1
// splitters at j * p2cap (first-level splitter j at (j + 1) p2cap - 1).
// Empty buckets get no workgroup and no bucket, so the final buckets stay in
// array order.  One workgroup; writes the segment table (nb1 entries), the
// count / scatter work list and hdr = {work items, count-matrix cells, final
// buckets}.
constexpr int kL2Threads = 1024;

__global__ void __launch_bounds__(kL2Threads)
k_part_level2(const int64_t* counts1, int64_t blocks1, int nb1, int64_t n, int p1, int p2, int p2cap, int64_t target,
              int64_t block_elems, PartBlock* work, PartSeg* segs, int64_t* hdr) {
    const int stride = p2cap > 0 ? p2cap : p2;   // splitter-list entries per first-level part
    __shared__ int64_t sc[3][kL2Threads];
    const int tid = int(threadIdx.x);
    const int per = (nb1 + kL2Threads - 1) / kL2Threads;
    const int b0 = tid * per < nb1 ? tid * per : nb1;
    const int b1 = b0 + per < nb1 ? b0 + per : nb1;
    struct Geo {
        int64_t begin, end, blocks;
        int ns, nbs;
    };
    auto geo = [&](int b) {
        Geo g;
        g.begin = counts1[int64_t(b) * blocks1];
        g.end = b + 1 < nb1 ? counts1[int64_t(b + 1) * blocks1] : n;
        const int64_t len = g.end - g.begin;
        if (p2cap > 0) g.ns = (b & 1) ? 0 : part_p2(len, target, p2cap) - 1;   // own sample: own part count
        else g.ns = ((b & 1) || len <= target) ? 0 : p2 - 1;
        g.blocks = len > 0 ? (len + block_elems - 1) / block_elems : 0;
        g.nbs = len > 0 ? 2 * g.ns + 1 : 0;
        return g;
    };
    int64_t nblk = 0, nbkt = 0, ncel = 0;
    for (int b = b0; b < b1; ++b) {
        const Geo g = geo(b);
        nblk += g.blocks;
        nbkt += g.nbs;
        ncel += g.blocks * g.nbs;
    }
    sc[0][tid] = nblk;
    sc[1][tid] = nbkt;
    sc[2][tid] = ncel;
    __syncthreads();
    for (int w = 1; w < kL2Threads; w <<= 1) {   // inclusive scans
        const int64_t a0 = tid >= w ? sc[0][tid - w] : 0;
        const int64_t a1 = tid >= w ? sc[1][tid - w] : 0;
        const int64_t a2 = tid >= w ? sc[2][tid - w] : 0;
        __syncthreads();
        sc[0][tid] += a0;
        sc[1][tid] += a1;
        sc[2][tid] += a2;
        __syncthreads();
    }
    int64_t blk = sc[0][tid] - nblk, bkt = sc[1][tid] - nbkt, cel = sc[2][tid] - ncel;
    for (int b = b0; b < b1; ++b) {
        const Geo g = geo(b);
        const int j = b / 2;
        const bool split = g.ns > 0;
        segs[b] = PartSeg{cel, bkt, int32_t(g.blocks), g.nbs, (b & 1) ? 1 : 0, 0};
        for (int64_t k = 0; k < g.blocks; ++k) {
            const int64_t x0 = g.begin + k * block_elems;
            work[blk + k] = PartBlock{x0, x0 + block_elems < g.end ? x0 + block_elems : g.end, cel + k,
                                      int32_t(g.blocks), split ? j * stride : 0, g.ns,
                                      split && j > 0 ? j * stride - 1 : -1,
                                      split && j < p1 - 1 ? (j + 1) * stride - 1 : -1, 0};
        }
        blk += g.blocks;
        bkt += g.nbs;
        cel += g.blocks * g.nbs;
    }
    if (tid == kL2Threads - 1) {
        hdr[0] = blk;
        hdr[1] = cel;
        hdr[2] = bkt;
    }
}

// ---------------------------------------------------------------- device plan
// The final ragged sort is planned on the device from the last level's
// bucket starts: every bucket is a segment, classed by ceil(log2 length);
// the list holds the bucket ids grouped by class, largest class first, and
// work items cover `spi` consecutive list entries of one class, as the host
// planner lays them out (the order inside a class is free: every segment is
// sorted on its own).  Equality buckets (one key value each) are left as
// they are when the sort is in place.  A longer bucket than the chunk gets
// class kPlanLong (chunk sorts and merge passes).  A bucket a little longer
// than the single-wave tile, 2^split < len <= 2^(split+1), is listed twice,
// as its head in class `split` and its tail in the tail's own class, and
// goes to the split list, whose buckets k_part_split_merge merges after the
// persistent sort: two single-wave tiles and a merge cost less than one
// padded tile over several waves (cls = kPlanSplitCls + tail class).
constexpr int kPlanClasses = 16;
constexpr int kPlanLong = 14;
constexpr int kPlanSplit = 15;       // ccount slot: split buckets
constexpr int kPlanSplitCls = 16;

struct PlanTable {
    int maxb;
    int small_max;
    int split;   // log2 of the head of a split bucket; -1: no splitting
    int spi[kPlanClasses];
    int variant[kPlanClasses];
};

// meta: ccount[16], cursor[16], (unused)[32], hdr[4], then the persistent
// kernels' two item counters
constexpr int kPlanMeta = 4 * kPlanClasses + 4;


// Adds one to h[c] for every lane of the wave with 0 <= c < kPlanClasses,
// c != kPlanLong, and
// returns the lane's previous count (its slot among the class's entries):
// one LDS atomic per distinct class of the wave (the buckets of a level
// mostly share one or two classes; per-lane atomics on one address
// serialise).  Every lane of the wave calls it.
__device__ __forceinline__ int plan_count(int* h, int c) {
    const int lane = int(__lane_id());
    uint64_t todo = __ballot(c >= 0 && c < kPlanClasses && c != kPlanLong);
    int slot = 0;
    while (todo) {
        const int first = __ffsll(static_cast<unsigned long long>(todo)) - 1;
        const int cf = __builtin_amdgcn_readlane(c, first);
        const uint64_t same = __ballot(c == cf) & todo;
        int base = 0;
        if (lane == first) base = atomicAdd(&h[cf], __popcll(same));
        base = __builtin_amdgcn_readlane(base, first);
        if (c == cf) slot = base + __popcll(same & ((uint64_t(1) << lane) - 1));
        todo &= ~same;
    }
    return slot;
}

// Class of final bucket i: c1 = the class it is listed in (a split bucket:
// its head's), c2 = a split bucket's tail class (else -1); cls[i] as above.
struct PlanClass {
    int cls, c1, c2;
};

__device__ __forceinline__ PlanClass plan_class(int64_t len, bool eq, int64_t chunk, int split) {
    const int64_t head = split >= 0 ? int64_t(1) << split : 0;
    PlanClass p{255, -1, -1};   // empty, or an equality bucket sorted in place: nothing to do
    if (eq) return p;
    if (len > chunk) {
        p.cls = kPlanLong;
    } else if (split >= 0 && len > head && len <= 2 * head) {
        const int64_t tl = len - head;
        p.c2 = tl > 1 ? 64 - __clzll(uint64_t(tl - 1)) : 0;
        p.c1 = split;
        p.cls = kPlanSplitCls + p.c2;
    } else if (len > 0) {
        p.cls = len > 1 ? 64 - __clzll(uint64_t(len - 1)) : 0;
        p.c1 = p.cls;
    }
    return p;
}

// Few workgroups, each over a stride of buckets, so that every class count
// reaches device memory with one atomic per workgroup (same-address device
// atomics serialise: one per bucket workgroup of 256 cost ~35 us at 2^28).
constexpr int kPlanThreads = 1024;
constexpr int kPlanGrid = 64;

__global__ void __launch_bounds__(kPlanThreads)
k_plan_classify(int64_t* starts, const unsigned char* equal, int skip_equal, int64_t nb, const int64_t* d_nb,
                int64_t n, int64_t chunk, int split, unsigned char* cls, int* ccount, int* long_list) {
    __shared__ int h[kPlanClasses];
    if (threadIdx.x < kPlanClasses) h[threadIdx.x] = 0;
    __syncthreads();
    if (d_nb) nb = *d_nb;   // bucket count of a level planned on the device
    const int64_t stride = int64_t(gridDim.x) * kPlanThreads;
    for (int64_t i0 = int64_t(blockIdx.x) * kPlanThreads; i0 < nb; i0 += stride) {   // uniform trip count
        const int64_t i = i0 + threadIdx.x;
        PlanClass p{255, -1, -1};
        if (i < nb) {
            const int64_t len = (i + 1 < nb ? starts[i + 1] : n) - starts[i];
            p = plan_class(len, skip_equal && equal[i], chunk, split);
            cls[i] = static_cast<unsigned char>(p.cls);
            if (p.cls == kPlanLong) long_list[atomicAdd(&ccount[kPlanLong], 1)] = int(i);   // rare
        }
        plan_count(h, p.c1);
        plan_count(h, p.c2 >= 0 ? kPlanSplit : -1);   // split buckets, counted in slot kPlanSplit
        plan_count(h, p.c2);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) starts[nb] = n;   // closing offset of the segment CSR
    __syncthreads();
    if (threadIdx.x < kPlanClasses && threadIdx.x != kPlanLong && h[threadIdx.x])
        atomicAdd(&ccount[threadIdx.x], h[threadIdx.x]);
}

// Layout of the class-grouped list and the item ranges, classes from maxb
// down: cstart[c] = first list entry of class c, istart[c] = its first
// work item.  Every workgroup of k_plan_fill / k_plan_items recomputes it
// (sixteen classes) from the class counts instead of a launch of its own.
struct PlanLayout {
    int cstart[kPlanClasses];
    int istart[kPlanClasses];
    int split;   // work items of the multi-wave kernel
    int items;   // all work items
};

__device__ void plan_layout(const int* ccount, const PlanTable& t, PlanLayout* L) {
    int seg = 0, items = 0, split = 0;
    for (int c = t.maxb; c >= 0; --c) {
        const int k = (ccount[c] + t.spi[c] - 1) / t.spi[c];
        L->cstart[c] = seg;
        L->istart[c] = items;
        seg += ccount[c];
        items += k;
        if (c > t.small_max) split += k;
    }
    L->split = split;
    L->items = items;
}

// Two passes over the workgroup's stride of buckets: the class counts (one
// device atomic per class for the workgroup's places in the list and in
// the split list), then the entries, in the same order.
__global__ void __launch_bounds__(kPlanThreads)
k_plan_fill(const unsigned char* cls, int64_t nb, const int64_t* d_nb, int* meta, PlanTable t, int* list,
            int* split_list) {
    __shared__ int h[kPlanClasses], base[kPlanClasses];
    __shared__ PlanLayout L;
    int* cursor = meta + kPlanClasses;
    if (threadIdx.x < kPlanClasses) h[threadIdx.x] = 0;
    if (threadIdx.x == 0) plan_layout(meta, t, &L);
    if (d_nb) nb = *d_nb;
    __syncthreads();
    const int64_t stride = int64_t(gridDim.x) * kPlanThreads;
    auto classes = [&](int64_t i, int& c1, int& c2) {
        const int c = i < nb ? int(cls[i]) : 255;
        const bool split = c >= kPlanSplitCls && c != 255;
        c1 = c < kPlanLong ? c : (split ? t.split : -1);   // the segment, or a split bucket's head
        c2 = split ? c - kPlanSplitCls : -1;               // a split bucket's tail
    };
    for (int64_t i0 = int64_t(blockIdx.x) * kPlanThreads; i0 < nb; i0 += stride) {
        int c1, c2;
        classes(i0 + threadIdx.x, c1, c2);
        plan_count(h, c1);
        plan_count(h, c2 >= 0 ? kPlanSplit : -1);
        plan_count(h, c2);
    }
    __syncthreads();
    if (threadIdx.x < kPlanClasses && threadIdx.x != kPlanLong) {
        base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], h[threadIdx.x]) : 0;
        h[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int64_t i0 = int64_t(blockIdx.x) * kPlanThreads; i0 < nb; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        int c1, c2;
        classes(i, c1, c2);
        const int l1 = plan_count(h, c1);
        const int ls = plan_count(h, c2 >= 0 ? kPlanSplit : -1);
        const int l2 = plan_count(h, c2);
        if (c1 >= 0) list[L.cstart[c1] + base[c1] + l1] = int(i) | (c2 >= 0 ? kSegHead : 0);
        if (c2 >= 0) {
            list[L.cstart[c2] + base[c2] + l2] = int(i) | kSegTail;
            split_list[base[kPlanSplit] + ls] = int(i);
        }
    }
}

// Work items of the class-grouped list; workgroup 0 also writes the item
// ranges the persistent kernels read: hdr = {multi-wave items, all items}.
__global__ void k_plan_items(int* meta, PlanTable t, SegItem* items) {
    __shared__ PlanLayout L;
    const int* ccount = meta;
    int* hdr = meta + 4 * kPlanClasses;
    if (threadIdx.x == 0) plan_layout(ccount, t, &L);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        hdr[0] = L.split;
        hdr[1] = L.items;
    }
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= L.items) return;
    int c = t.maxb;   // classes run from maxb down, their item ranges upward
    while (c > 0 && k >= L.istart[c] + (ccount[c] + t.spi[c] - 1) / t.spi[c]) --c;
    const int j = k - L.istart[c];
    SegItem it;
    it.bucket = c;
    it.variant = t.variant[c];
    it.first = L.cstart[c] + j * t.spi[c];
    it.count = min(t.spi[c], ccount[c] - j * t.spi[c]);
    items[k] = it;
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

// n bounds the count; with d_n the count is *d_n (known on the device), and
// the workgroups past it return at once.
hipError_t exclusive_scan(int64_t* v, int64_t n, int64_t* sums, hipStream_t s, const int64_t* d_n = nullptr) {
    const int64_t blocks = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_sums, dim3(unsigned(blocks)), dim3(kScanThreads), 0, s, v, n, sums, d_n);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, sums, blocks);
    hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(blocks)), dim3(kScanThreads), 0, s, v, n, sums, d_n);
    return hipGetLastError();
}

}  // namespace

namespace {

// Elements per workgroup of count / scatter: 512 KiB of elements while that
// leaves at least 1024 workgroups, else halved down to 64 KiB (fewer
// workgroups leave CUs idle at the grid's end; fewer and longer ones shrink
// the count matrix and its scan).  Elements per workgroup, ms per sort:
// 2^28 int32 32768 3.94, 65536 3.60, 131072 3.50, 262144 3.81; 2^26 int32
// 65536 1.06, 131072 1.16; 2^27 uint64 32768 4.06, 65536 4.03, 131072 4.08;
// 2^25 int64+int32 records 16384 2.81, 32768 2.83, 65536 2.85.
static int64_t part_block_elems(size_t esz, int64_t n) {
    int64_t e = int64_t(512 * 1024) / int64_t(esz);
    const int64_t floor_e = int64_t(64 * 1024) / int64_t(esz);
    while (e > floor_e && n / e < 1024) e /= 2;
    return e;
}

// Most first-level workgroups of any sort of up to n elements (a smaller
// sort can take smaller workgroups, and so more of them: fewer than 2048
// whenever it does).
static int64_t part_blocks_bound(size_t esz, int64_t n) {
    const int64_t e0 = int64_t(512 * 1024) / int64_t(esz), floor_e = int64_t(64 * 1024) / int64_t(esz);
    return std::max((n + e0 - 1) / e0, std::min<int64_t>(2048, (n + floor_e - 1) / floor_e));
}
constexpr int kPartOneLevel = 256;           // parts up to which one level is used

// Parts of a partition sort of n elements (the caller's MinNbParts /
// MaxNbParts bound them when nonzero): p1 first-level x p2 second-level
// parts (p2 = 1: one level) and m sampled keys.
struct PartShape {
    int64_t target = 0;
    int p1 = 0, p2 = 1;
    int64_t parts = 0, m = 0;
    // two-stage sampling (two levels): m1 = s1 * p1 keys for the first
    // level only, then s2 keys per second-level part from each first-level
    // bucket, which gets part_p2(len, target2, p2cap) parts (p2cap * s2 keys
    // fit one chunk; target2 folds a caller's part bound into the bucket
    // size)
    bool two_stage = false;
    int s1 = 0, s2 = 0, p2cap = 0;
    int64_t m1 = 0, target2 = 0;
};

PartShape part_shape(int dtype, int64_t n, int min_parts, int max_parts) {
    const size_t esz = cs_dtype_size(dtype);
    PartShape sh;
    // Buckets average 1536 elements, 3/4 of a 2048-key tile: with 12-64
    // samples per part most fill that power of two to ~75 %, and the largest
    // of ~10^5 buckets stays far under the chunk (8192 4-byte / 4096 wider
    // elements) above which a bucket takes the long path.  2^28 int32 (ms
    // per sort, 16 samples per part): 1400 3.80, 1536 3.76, 1700 3.91,
    // 3072 4.13; the 2048-key tile sorts a key in fewer layers than the
    // 4096-key one, which outweighs twice the parts.  With split buckets and
    // two-stage sampling the optimum is flat: 1280 3.42, 1408 3.40, 1536
    // 3.38-3.39, 1664 3.40, 1792 3.40, 2048 3.42.
    sh.target = 1536;
    int64_t P = (n + sh.target - 1) / sh.target;
    if (min_parts > 0) P = std::max<int64_t>(P, min_parts);
    if (max_parts > 0) P = std::min<int64_t>(P, max_parts);
    const int maxp = part_kernels_for(dtype).max_splitters + 1;   // parts per level
    const int64_t pmax = int64_t(maxp) * maxp;
    P = std::max<int64_t>(2, std::min(P, pmax));
    sh.p1 = int(std::min<int64_t>(P, maxp));
    if (P > kPartOneLevel) {
        sh.p1 = int(std::ceil(std::sqrt(double(P))));
        sh.p2 = int(P / sh.p1);   // p1 * p2 <= P: a caller's part bound holds
        sh.p1 = std::min(sh.p1, maxp);
        sh.p2 = std::min(sh.p2, maxp);
    }
    sh.parts = int64_t(sh.p1) * sh.p2;
    // Samples per part, two levels.  2^28 int32 (ms per sort): 8 per part
    // 3.77, 12 3.75, 16 3.76, 24 3.78 (a smaller sample to sort against a
    // wider spread of bucket sizes); 2^27 uint64: 16 4.42, 32 4.25, 48 4.23
    // (8-byte keys make the bucket sort dearer per spilled key), so wider
    // elements keep 32.
    sh.m = sh.parts * (sh.p2 > 1 ? (esz == 4 ? 12 : 32) : 64);
    // Two-stage sampling for two levels: 16 keys per first-level part, 12
    // per second-level part from each first-level bucket.  2^28 int32 (ms
    // per sort): one global sample 3.50; two-stage with p2 fixed and 16 / 32
    // / 64 first-level and 12 second-level samples per part 3.42 / 3.43 /
    // 3.48, with 8 / 16 second-level 3.45 / 3.43; 2^26 int32 1.06 -> 0.99,
    // 2^26 int32+int32 records 1.96 -> 1.88, 2^27 uint64 4.03 -> 3.92; 2^28
    // int32 in presorted runs 3.70 either way.  The global sample's merge
    // sort (eight passes of small, latency-bound kernels) gives way to one
    // chunk sort of all parts' samples.  Each first-level bucket then takes
    // as many second-level parts as its own size needs (the small
    // first-level sample leaves those sizes ~25 % apart).
    const LargeKernels lk = large_kernels_for(dtype);
    sh.s1 = 16;
    sh.s2 = 12;   // 2^27 uint64: 12 / 16 / 20 per part 3.87 / 3.92 / 3.98 ms
    sh.p2cap = std::min(maxp, lk.chunk / sh.s2);
    sh.two_stage = sh.p2 > 1 && sh.p2cap > 1;
    if (sh.two_stage && sh.p2 > sh.p2cap * 3 / 4) {
        // A first-level bucket needs up to ~1.3x the average part count (16
        // samples per part spread their sizes by ~25 %): more first-level
        // parts keep the average at 3/4 of the cap (8-byte keys: 341 parts
        // at most, so that 12 samples per part fit a 4096-key chunk).
        sh.p2 = sh.p2cap * 3 / 4;
        sh.p1 = int(std::min<int64_t>(maxp, (P + sh.p2 - 1) / sh.p2));
        sh.p2 = std::min<int64_t>(sh.p2, P / sh.p1);   // p1 * p2 <= P: a caller's part bound holds
        sh.parts = int64_t(sh.p1) * sh.p2;
        sh.m = sh.parts * (esz == 4 ? 12 : 32);
    }
    sh.m1 = int64_t(sh.s1) * sh.p1;
    sh.target2 = std::max<int64_t>(1, (n + sh.parts - 1) / sh.parts);
    return sh;
}

// Second-level sizes: splitter-list entries per first-level part, buckets
// per segment at most, final buckets at most and count-matrix cells at most
// (the grids and temporaries are sized by them; the device plan holds the
// actual counts).
struct Level2Bounds {
    int stride = 0;
    int64_t nbs_max = 0, fnb = 0, cells = 0;
};

Level2Bounds level2_bounds(const PartShape& sh, int64_t n, int64_t blocks1) {
    Level2Bounds b;
    const int64_t nb1 = 2 * int64_t(sh.p1) - 1;
    b.stride = sh.two_stage ? sh.p2cap : sh.p2;
    b.nbs_max = 2 * int64_t(b.stride) - 1;
    b.fnb = b.nbs_max * nb1;
    // two-stage: the parts' p2 sum to at most n / target2 + p1, each part
    // has 2 p2 - 1 buckets, each equality bucket one
    if (sh.two_stage) b.fnb = std::min(b.fnb, 2 * ((n + sh.target2 - 1) / sh.target2) + nb1);
    b.cells = b.nbs_max * (blocks1 + nb1);   // every segment rounds its blocks up once
    return b;
}

}  // namespace

// Sorts data[0, n) into out (out may equal data) with `values` as scratch of
// at least n elements and the call's other temporaries from `tmp`.  Sample
// sort into buckets of 1536 elements: one partition level (data -> values)
// for up to kPartOneLevel parts, else two (data -> values -> out) with p1 x
// p2 parts in one splitter list: the first-level splitters from a small
// sample, each first-level bucket's second-level splitters from its own
// sample (or, where that sample would not fit a chunk, all from one global
// sample).  The buckets are then sorted as one ragged batch in place in
// out, planned on the device; buckets just over the single-wave tile are
// sorted as head and tail and merged, and the rare bucket longer than a
// chunk by chunk sorts and merge passes.  Everything is laid out on the
// device (the second level's splitters and work list, the final plan, the
// long buckets' chunks and tiles): the call enqueues its work on the stream
// and returns without waiting.
int partition_sort(int dtype, void* data, int64_t n, void* out, void* values, size_t values_elems, int min_parts,
                   int max_parts, Scratch& tmp) {
    const LargeKernels lk = large_kernels_for(dtype);
    const PartKernels pk = part_kernels_for(dtype);
    const size_t esz = cs_dtype_size(dtype);
    const hipStream_t stream = tmp.stream();
    if (values_elems < size_t(n)) return CS_E_WORKSPACE;
    if (n <= lk.chunk) return cs_batched_sort(dtype, data, out, 1, int32_t(n), CS_METHOD_AUTO, 0, 0, stream);

    const PartShape sh = part_shape(dtype, n, min_parts, max_parts);
    const int p1 = sh.p1, p2 = sh.p2;
    const int64_t parts = sh.parts, m = sh.m, target = sh.target;
    const int64_t block_elems = part_block_elems(esz, n);
    const int64_t blocks1 = (n + block_elems - 1) / block_elems;
    const int nb1 = 2 * p1 - 1;
    const int64_t cells1 = int64_t(nb1) * blocks1;
    // long-bucket sort (final buckets longer than the tile)
    const int64_t nlong_max = n / lk.chunk + 1;
    const int64_t ntile_max = n / lk.tile + nlong_max;

    unsigned char *samples = nullptr, *samples_tmp = nullptr, *gsp = nullptr, *l1sp = nullptr, *smp2 = nullptr;
    int64_t* cor = nullptr;
    PartBlock* work1 = nullptr;
    PartSeg* segs1 = nullptr;
    int64_t *counts1 = nullptr, *sums1 = nullptr;
    const int64_t msort = sh.two_stage ? sh.m1 : m;   // keys sampled and sorted before the first level
    int rc = tmp.get(&samples, size_t(msort) * esz);
    if (rc == CS_OK) rc = tmp.get(&samples_tmp, size_t(msort) * esz);
    if (rc == CS_OK && sh.two_stage) rc = tmp.get(&smp2, size_t(p1) * size_t(lk.chunk) * esz);
    const int64_t nsplitters = sh.two_stage ? int64_t(p1) * sh.p2cap - 1 : parts - 1;
    if (rc == CS_OK) rc = tmp.get(&gsp, size_t(nsplitters) * size_t(pk.key_size));
    if (rc == CS_OK && p2 > 1) rc = tmp.get(&l1sp, size_t(std::max(p1 - 1, 1)) * size_t(pk.key_size));
    if (rc == CS_OK) rc = tmp.get(&cor, size_t(ntile_max) + 1);
    if (rc == CS_OK) rc = tmp.get(&work1, size_t(blocks1));
    if (rc == CS_OK) rc = tmp.get(&segs1, 1);
    if (rc == CS_OK) rc = tmp.get(&counts1, size_t(cells1));
    if (rc == CS_OK) rc = tmp.get(&sums1, size_t((cells1 + kScanTile - 1) / kScanTile) + 1);
    if (rc != CS_OK) return rc;

    // hashed sample sorted by the merge path; splitters and the first
    // level's work list from it
    hipError_t e = pk.sample(data, n, samples, int(msort), stream);
    if (e != hipSuccess) return int(e);
    rc = merge_sort(dtype, samples, msort, samples, samples_tmp, tmp);
    if (rc != CS_OK) return rc;
    if (sh.two_stage)   // first-level splitters only (one level of p1 parts from the small sample)
        e = pk.prep(samples, int(msort), p1, p1, 1, l1sp, nullptr, n, block_elems, blocks1, work1, segs1, stream);
    else
        e = pk.prep(samples, int(m), int(parts), p1, p2, gsp, l1sp, n, block_elems, blocks1, work1, segs1, stream);
    if (e != hipSuccess) return int(e);

    // level 1: data -> values (straight into out for one level of an
    // out-of-place sort), one segment over the whole array
    const void* sp1 = p2 > 1 ? static_cast<const void*>(l1sp) : static_cast<const void*>(gsp);
    void* dst1 = (p2 == 1 && out != data) ? out : values;
    e = pk.count(data, sp1, work1, int(blocks1), p1 - 1, counts1, nullptr, stream);
    if (e == hipSuccess) e = exclusive_scan(counts1, cells1, sums1, stream);
    if (e == hipSuccess) e = pk.scatter(data, sp1, work1, int(blocks1), p1 - 1, counts1, dst1, nullptr, stream);
    // one level in place: the buckets go back into out, so that the final
    // sort also runs in place (equality buckets then stay as they are)
    if (e != hipSuccess) return int(e);
    if (p2 == 1 && dst1 != out) {
        rc = cs_device_copy(out, values, size_t(n) * esz, stream);
        if (rc != CS_OK) return rc;
    }
    // two-stage sampling: the second-level splitters from each first-level
    // bucket's own sample (counts1 now holds the bucket starts)
    if (sh.two_stage) {
        e = pk.l2_gather(values, counts1, blocks1, p1, sh.target2, sh.p2cap, sh.s2, lk.chunk, n, smp2, stream);
        if (e == hipSuccess) e = lk.chunk_sort(smp2, smp2, int64_t(p1) * lk.chunk, stream);
        if (e == hipSuccess)
            e = pk.l2_pick(smp2, counts1, blocks1, p1, sh.target2, sh.p2cap, sh.s2, lk.chunk, n, l1sp, gsp, stream);
        if (e != hipSuccess) return int(e);
    }

    // final plan's counters (meta, then the persistent kernels' two), cleared
    // by the last level's bucket-starts kernel
    constexpr int kZero = kPlanMeta + 2;
    int* meta = nullptr;
    rc = tmp.get(&meta, kZero);
    if (rc != CS_OK) return rc;
    int* d_counters = meta + kPlanMeta;
    int64_t* d_fst = nullptr;
    unsigned char* d_feq = nullptr;
    const int64_t* d_fnb = nullptr;   // final bucket count on the device (two levels)
    int64_t fnb = nb1;                // its bound
    if (p2 == 1) {
        rc = tmp.get(&d_fst, size_t(nb1) + 1);
        if (rc == CS_OK) rc = tmp.get(&d_feq, size_t(nb1));
        if (rc != CS_OK) return rc;
        hipLaunchKernelGGL(k_part_starts, dim3(1), dim3(256), 0, stream, segs1, counts1, d_fst, d_feq, meta, kZero);
        e = hipGetLastError();
    } else {
        // level 2: values -> out, one segment per first-level bucket, laid
        // out on the device; grids are sized by the bounds and the surplus
        // workgroups return at once
        const Level2Bounds l2 = level2_bounds(sh, n, blocks1);
        const int64_t blocks2 = blocks1 + nb1;   // every segment rounds up once
        const int64_t cells2 = l2.cells;
        fnb = l2.fnb;
        PartBlock* work2 = nullptr;
        PartSeg* segs2 = nullptr;
        int64_t *hdr2 = nullptr, *counts2 = nullptr, *sums2 = nullptr;
        rc = tmp.get(&work2, size_t(blocks2));
        if (rc == CS_OK) rc = tmp.get(&segs2, size_t(nb1));
        if (rc == CS_OK) rc = tmp.get(&hdr2, 4);
        if (rc == CS_OK) rc = tmp.get(&counts2, size_t(cells2));
        if (rc == CS_OK) rc = tmp.get(&sums2, size_t((cells2 + kScanTile - 1) / kScanTile) + 1);
        if (rc == CS_OK) rc = tmp.get(&d_fst, size_t(fnb) + 1);
        if (rc == CS_OK) rc = tmp.get(&d_feq, size_t(fnb));
        if (rc != CS_OK) return rc;
        hipLaunchKernelGGL(k_part_level2, dim3(1), dim3(kL2Threads), 0, stream, counts1, blocks1, nb1, n, p1, p2,
                           sh.two_stage ? sh.p2cap : 0, sh.two_stage ? sh.target2 : target, block_elems, work2,
                           segs2, hdr2);
        e = hipGetLastError();
        // the scan covers the hdr2[1] cells the level uses
        if (e == hipSuccess) e = pk.count(values, gsp, work2, int(blocks2), l2.stride - 1, counts2, hdr2, stream);
        if (e == hipSuccess) e = exclusive_scan(counts2, cells2, sums2, stream, hdr2 + 1);
        if (e == hipSuccess)
            e = pk.scatter(values, gsp, work2, int(blocks2), l2.stride - 1, counts2, out, hdr2, stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_part_starts, dim3(unsigned(nb1)), dim3(256), 0, stream, segs2, counts2, d_fst, d_feq,
                               meta, kZero);
            e = hipGetLastError();
        }
        d_fnb = hdr2 + 2;
    }
    if (e != hipSuccess) return int(e);

    // plan the final ragged sort on the device
    const SegClasses sc = seg_classes(dtype);
    PlanTable t{};
    t.maxb = sc.maxb;
    t.small_max = sc.small_max;
    t.split = sc.small_max < sc.maxb ? sc.small_max : -1;
    for (int c = 0; c < kPlanClasses; ++c) {
        t.spi[c] = c <= sc.maxb ? std::max(1, sc.spi[c]) : 1;
        t.variant[c] = c <= sc.maxb ? sc.variant[c] : 0;
    }
    unsigned char* cls = nullptr;
    int *d_list = nullptr, *long_list = nullptr, *split_list = nullptr;
    SegItem* d_items = nullptr;
    rc = tmp.get(&cls, size_t(fnb));
    if (rc == CS_OK) rc = tmp.get(&d_list, 2 * size_t(fnb));     // a split bucket is listed twice
    if (rc == CS_OK) rc = tmp.get(&d_items, 2 * size_t(fnb));
    if (rc == CS_OK) rc = tmp.get(&long_list, size_t(nlong_max));
    if (rc == CS_OK) rc = tmp.get(&split_list, size_t(fnb));
    if (rc != CS_OK) return rc;
    const unsigned g = unsigned((fnb + 255) / 256);
    // the final sort runs in place: equality buckets need no sort
    const unsigned gp = unsigned(std::min<int64_t>(kPlanGrid, (fnb + kPlanThreads - 1) / kPlanThreads));
    hipLaunchKernelGGL(k_plan_classify, dim3(gp), dim3(kPlanThreads), 0, stream, d_fst, d_feq, 1, fnb, d_fnb, n,
                       int64_t(lk.chunk), t.split, cls, meta, long_list);
    hipLaunchKernelGGL(k_plan_fill, dim3(gp), dim3(kPlanThreads), 0, stream, cls, fnb, d_fnb, meta, t, d_list,
                       split_list);
    hipLaunchKernelGGL(k_plan_items, dim3(2 * g), dim3(256), 0, stream, meta, t, d_items);   // up to 2 fnb items
    e = hipGetLastError();
    if (e != hipSuccess) return int(e);
    SegShape ss;
    ss.grid[kSegMulti] = seg_persistent_grid(dtype, kSegMulti, tmp.grid_cap);
    ss.grid[kSegSingle] = seg_persistent_grid(dtype, kSegSingle, tmp.grid_cap);
    if (ss.grid[kSegMulti] <= 0 || ss.grid[kSegSingle] <= 0) return CS_E_NO_DEVICE;
    e = seg_launch_all(dtype, ss, out, out, d_fst, d_list, d_items, 0, d_counters, stream, meta + 4 * kPlanClasses,
                       false);
    // the two sorted halves of every split bucket (meta[kPlanSplit] of them)
    if (e == hipSuccess && t.split >= 0)
        e = pk.split_merge(out, d_fst, split_list, meta + kPlanSplit, tmp.grid_cap, stream);
    // buckets longer than the tile (meta[kPlanLong] of them): chunk sorts and
    // merge passes between out and values
    if (e == hipSuccess)
        e = pk.long_sort(out, out, values, d_fst, long_list, meta + kPlanLong, cor, tmp.grid_cap, stream);
    return e == hipSuccess ? CS_OK : int(e);
}

// Every temporary partition_sort can take for n elements, at its largest:
// the data-dependent second level and the host-planned fallback are bounded
// by their part and chunk counts.
size_t partition_scratch_bytes(int dtype, int64_t n, int min_parts, int max_parts) {
    const LargeKernels lk = large_kernels_for(dtype);
    const PartKernels pk = part_kernels_for(dtype);
    const size_t esz = cs_dtype_size(dtype);
    if (!lk.chunk || n <= lk.chunk) return 0;
    const PartShape sh = part_shape(dtype, n, min_parts, max_parts);
    auto r = [](size_t bytes) { return scratch_round(bytes ? bytes : 1); };
    size_t total = 0;
    const int64_t nlong_max = n / lk.chunk + 1;
    const int64_t ntile_max = n / lk.tile + nlong_max;
    // Two levels: p1 = ceil(sqrt(P)) grows with n, while p2 = P / p1 and the
    // part count p1 * p2 do not always; sizing parts, samples and second-level
    // buckets by p1 * p1 keeps this bound growing with n, so a workspace
    // reserved for n holds every smaller sort.
    const int64_t parts = sh.p2 > 1 ? int64_t(sh.p1) * sh.p1 : sh.parts;
    const int64_t m = sh.m / sh.parts * parts;
    // samples, their sorted copy and merge sort: one global sample, or the
    // small first-level sample and a chunk per first-level part (the larger)
    const size_t global_smp = 2 * r(size_t(m) * esz) + merge_scratch_bytes(dtype, m);
    const size_t staged_smp = sh.p2 > 1 ? 2 * r(size_t(sh.m1) * esz) + merge_scratch_bytes(dtype, sh.m1) +
                                              r(size_t(sh.p1) * size_t(lk.chunk) * esz)
                                        : 0;
    total += std::max(global_smp, staged_smp);
    const int64_t nsplit_b = std::max<int64_t>(parts - 1, sh.p2 > 1 ? int64_t(sh.p1) * sh.p2cap - 1 : 0);
    total += r(size_t(nsplit_b) * size_t(pk.key_size));                   // splitter list
    total += r(size_t(std::max(sh.p1 - 1, 1)) * size_t(pk.key_size));     // first-level splitters
    total += r(size_t(ntile_max + 1) * 8) + r(size_t(nlong_max) * 4);     // long buckets: co-ranks, list
    total += r(size_t(kPlanMeta + 2) * 4);                                // plan counters
    auto level = [&](int64_t segs, int64_t blocks, int64_t cells) {
        size_t b = r(size_t(blocks) * sizeof(PartBlock)) + r(size_t(segs) * sizeof(PartSeg));
        b += r(size_t(cells) * 8) + r(size_t((cells + kScanTile - 1) / kScanTile + 1) * 8);
        return b;
    };
    const int64_t blocks1 = part_blocks_bound(esz, n);
    const int64_t nb1 = 2 * int64_t(sh.p1) - 1;
    total += level(1, blocks1, nb1 * blocks1);
    int64_t fnb = nb1;
    if (sh.p2 > 1) {
        // buckets per second-level segment: p2 <= p1 with one global list,
        // at most p2cap with two-stage sampling (whose parts sum to at most
        // p1 * p1 + p1)
        const int64_t nbs = sh.two_stage ? 2 * int64_t(sh.p2cap) - 1 : 2 * int64_t(sh.p1) - 1;
        const int64_t blocks2 = blocks1 + nb1;                            // each segment rounds up once
        fnb = nbs * nb1;
        if (sh.two_stage) fnb = std::min(fnb, 2 * (int64_t(sh.p1) * sh.p1 + sh.p1) + nb1);
        total += level(nb1, blocks2, nbs * blocks2) + r(4 * 8);           // + device header
    }
    total += r(size_t(fnb + 1) * 8) + r(size_t(fnb));                     // final starts, equality flags
    // device plan of the final sort: classes, list and items (a split bucket
    // twice), split list
    total += r(size_t(fnb)) + r(2 * size_t(fnb) * 4) + r(2 * size_t(fnb) * sizeof(SegItem)) + r(size_t(fnb) * 4);
    return total;
}

}  // namespace cs
