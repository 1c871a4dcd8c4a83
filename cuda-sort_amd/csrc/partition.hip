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

__global__ void k_part_starts(const PartSeg* segs, const int64_t* offsets, int64_t* starts, unsigned char* equal) {
    const PartSeg g = segs[blockIdx.x];
    for (int b = int(threadIdx.x); b < g.nb; b += int(blockDim.x)) {
        starts[g.first + b] = offsets[g.cell0 + int64_t(b) * g.stride];
        equal[g.first + b] = (g.equal || (b & 1)) ? 1 : 0;
    }
}

// ---------------------------------------------------------------- device plan
// The final ragged sort is planned on the device from the last level's
// bucket starts: every bucket is a segment, classed by ceil(log2 length);
// the list holds the bucket ids grouped by class, largest class first, and
// work items cover `spi` consecutive list entries of one class, as the host
// planner lays them out (the order inside a class is free: every segment is
// sorted on its own).  Equality buckets (one key value each) are left as
// they are when the sort is in place.  A longer bucket than the tile gets
// class kPlanLong and sends the call to the host planner, which cuts it
// into chunks.
constexpr int kPlanClasses = 16;
constexpr int kPlanLong = 14;

struct PlanTable {
    int maxb;
    int small_max;
    int spi[kPlanClasses];
    int variant[kPlanClasses];
};

// meta: ccount[16], cursor[16], cstart[16], istart[16], hdr[4]
constexpr int kPlanMeta = 4 * kPlanClasses + 4;

__global__ void k_plan_zero(int* meta) {
    if (threadIdx.x < kPlanMeta) meta[threadIdx.x] = 0;
}

__global__ void k_plan_classify(int64_t* starts, const unsigned char* equal, int skip_equal, int64_t nb, int64_t n,
                                int64_t chunk, unsigned char* cls, int* ccount) {
    __shared__ int h[kPlanClasses];
    if (threadIdx.x < kPlanClasses) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < nb) {
        const int64_t len = (i + 1 < nb ? starts[i + 1] : n) - starts[i];
        int c = 255;   // empty, or an equality bucket sorted in place: nothing to do
        if (skip_equal && equal[i]) c = 255;
        else if (len > chunk) c = kPlanLong;
        else if (len > 0) c = len > 1 ? 64 - __clzll(uint64_t(len - 1)) : 0;
        cls[i] = static_cast<unsigned char>(c);
        if (c < kPlanClasses) atomicAdd(&h[c], 1);
    }
    if (i == 0) starts[nb] = n;   // closing offset of the segment CSR
    __syncthreads();
    if (threadIdx.x < kPlanClasses && h[threadIdx.x]) atomicAdd(&ccount[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_plan_layout(int* meta, PlanTable t) {
    if (threadIdx.x != 0) return;
    const int* ccount = meta;
    int* cstart = meta + 2 * kPlanClasses;
    int* istart = meta + 3 * kPlanClasses;
    int* hdr = meta + 4 * kPlanClasses;
    int seg = 0, items = 0, split = 0;
    for (int c = t.maxb; c >= 0; --c) {
        const int k = (ccount[c] + t.spi[c] - 1) / t.spi[c];
        cstart[c] = seg;
        istart[c] = items;
        seg += ccount[c];
        items += k;
        if (c > t.small_max) split += k;
    }
    hdr[0] = split;
    hdr[1] = items;
    hdr[2] = ccount[kPlanLong];
}

__global__ void k_plan_fill(const unsigned char* cls, int64_t nb, int* meta, int* list) {
    __shared__ int h[kPlanClasses], base[kPlanClasses];
    int* cursor = meta + kPlanClasses;
    const int* cstart = meta + 2 * kPlanClasses;
    if (threadIdx.x < kPlanClasses) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int c = i < nb ? int(cls[i]) : 255;
    const int local = c < kPlanLong ? atomicAdd(&h[c], 1) : 0;
    __syncthreads();
    if (threadIdx.x < kPlanClasses && h[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], h[threadIdx.x]);
    __syncthreads();
    if (c < kPlanLong) list[cstart[c] + base[c] + local] = int(i);
}

__global__ void k_plan_items(const int* meta, PlanTable t, SegItem* items) {
    const int* ccount = meta;
    const int* cstart = meta + 2 * kPlanClasses;
    const int* istart = meta + 3 * kPlanClasses;
    const int* hdr = meta + 4 * kPlanClasses;
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= hdr[1]) return;
    int c = t.maxb;   // classes run from maxb down, their item ranges upward
    while (c > 0 && k >= istart[c] + (ccount[c] + t.spi[c] - 1) / t.spi[c]) --c;
    const int j = k - istart[c];
    SegItem it;
    it.bucket = c;
    it.variant = t.variant[c];
    it.first = cstart[c] + j * t.spi[c];
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

hipError_t exclusive_scan(int64_t* v, int64_t n, int64_t* sums, hipStream_t s) {
    const int64_t blocks = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_sums, dim3(unsigned(blocks)), dim3(kScanThreads), 0, s, v, n, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, sums, blocks);
    hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(blocks)), dim3(kScanThreads), 0, s, v, n, sums);
    return hipGetLastError();
}

}  // namespace

namespace {

constexpr int64_t kPartBlockElems = 65536;   // elements per workgroup of count / scatter
constexpr int kPartOneLevel = 256;           // parts up to which one level is used

// One partition level: the workgroups of every segment, the segments' bucket
// start descriptors, and per segment its end and whether its buckets are
// equality buckets by construction (a first-level equality bucket).
struct Level {
    std::vector<PartBlock> work;
    std::vector<PartSeg> segs;
    std::vector<int64_t> seg_end;
    std::vector<char> seg_equal;
    int64_t cells = 0;
    int64_t nbuckets = 0;
    int max_ns = 0;
};

void clear_level(Level& L) {
    L.work.clear();
    L.segs.clear();
    L.seg_end.clear();
    L.seg_equal.clear();
    L.cells = 0;
    L.nbuckets = 0;
    L.max_ns = 0;
}

// Host arrays of one call, kept per thread so that their pages stay mapped
// from call to call (first-touch page faults of fresh multi-MB vectors cost
// more than the loops that fill them).
struct PartHost {
    Level l1, l2;
    std::vector<int64_t> st1, st2, seg;
    std::vector<MergeTile> tiles;
    std::vector<int> list;
    std::vector<SegItem> items;
};

void add_segment(Level& L, int64_t begin, int64_t end, int sp0, int ns, int lo, int hi, bool equal) {
    const int64_t len = end - begin;
    if (len <= 0) return;
    const int64_t blocks = (len + kPartBlockElems - 1) / kPartBlockElems;
    const int nb = 2 * ns + 1;
    L.segs.push_back(PartSeg{L.cells, L.nbuckets, int32_t(blocks), nb, equal ? 1 : 0, 0});
    L.seg_end.push_back(end);
    L.seg_equal.push_back(equal ? 1 : 0);
    for (int64_t k = 0; k < blocks; ++k) {
        const int64_t b0 = begin + k * kPartBlockElems;
        L.work.push_back(PartBlock{b0, std::min(end, b0 + kPartBlockElems), L.cells + k, int32_t(blocks), sp0, ns,
                                   lo, hi, 0});
    }
    L.cells += int64_t(nb) * blocks;
    L.nbuckets += nb;
    L.max_ns = std::max(L.max_ns, ns);
}

// Count, scan and scatter of one level (src -> dst) and its bucket starts in
// d_starts (nbuckets + 1 entries); with h_starts they are also read back
// (one host synchronisation).
int run_level(const PartKernels& pk, const Level& L, const void* src, const void* splitters, void* dst,
              int64_t* d_starts, unsigned char* d_equal, std::vector<int64_t>* h_starts, Scratch& tmp) {
    const hipStream_t stream = tmp.stream();
    PartBlock* work = nullptr;
    PartSeg* segs = nullptr;
    int64_t *counts = nullptr, *sums = nullptr;
    int rc = tmp.get(&work, L.work.size());
    if (rc == CS_OK) rc = tmp.get(&segs, L.segs.size());
    if (rc == CS_OK) rc = tmp.get(&counts, size_t(L.cells));
    if (rc == CS_OK) rc = tmp.get(&sums, size_t((L.cells + kScanTile - 1) / kScanTile) + 1);
    if (rc != CS_OK) return rc;
    hipError_t e = hipMemcpyAsync(work, L.work.data(), L.work.size() * sizeof(PartBlock), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(segs, L.segs.data(), L.segs.size() * sizeof(PartSeg), hipMemcpyHostToDevice, stream);
    const int blocks = int(L.work.size());
    if (e == hipSuccess) e = pk.count(src, splitters, work, blocks, L.max_ns, counts, stream);
    if (e == hipSuccess) e = exclusive_scan(counts, L.cells, sums, stream);
    if (e == hipSuccess) e = pk.scatter(src, splitters, work, blocks, L.max_ns, counts, dst, stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_part_starts, dim3(unsigned(L.segs.size())), dim3(256), 0, stream, segs, counts, d_starts,
                           d_equal);
        e = hipGetLastError();
    }
    if (h_starts) {
        h_starts->assign(size_t(L.nbuckets), 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_starts->data(), d_starts, h_starts->size() * sizeof(int64_t), hipMemcpyDeviceToHost,
                               stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
    }
    return e == hipSuccess ? CS_OK : int(e);
}

// Parts of a partition sort of n elements (the caller's MinNbParts /
// MaxNbParts bound them when nonzero): p1 first-level x p2 second-level
// parts (p2 = 1: one level) and m sampled keys.
struct PartShape {
    int64_t target = 0;
    int p1 = 0, p2 = 1;
    int64_t parts = 0, m = 0;
};

PartShape part_shape(int dtype, int64_t n, int min_parts, int max_parts) {
    const LargeKernels lk = large_kernels_for(dtype);
    const size_t esz = cs_dtype_size(dtype);
    PartShape sh;
    // Buckets average 3/8 of the tile (3072 4-byte / 1536 wider keys): with
    // 16-64 samples per part the largest of ~10^5 buckets stays under the
    // tile and most fill the next power of two to ~75 %.
    sh.target = lk.chunk * 3 / 8;
    int64_t P = (n + sh.target - 1) / sh.target;
    if (min_parts > 0) P = std::max<int64_t>(P, min_parts);
    if (max_parts > 0) P = std::min<int64_t>(P, max_parts);
    const int64_t pmax = int64_t(kPartMaxSplitters + 1) * (kPartMaxSplitters + 1);
    P = std::max<int64_t>(2, std::min(P, pmax));
    sh.p1 = int(P);
    if (P > kPartOneLevel) {
        sh.p1 = int(std::ceil(std::sqrt(double(P))));
        sh.p2 = int(P / sh.p1);   // p1 * p2 <= P: a caller's part bound holds
        sh.p1 = std::min(sh.p1, kPartMaxSplitters + 1);
        sh.p2 = std::min(sh.p2, kPartMaxSplitters + 1);
    }
    sh.parts = int64_t(sh.p1) * sh.p2;
    // Samples per part.  Two levels, 2^28 int32: 16 per part 4.97 ms, 32 per
    // part 5.09 ms (a smaller sample to sort outweighs a wider spread of
    // bucket sizes); 2^27 uint64: 5.20 ms against 4.92 (its buckets of 1536
    // spill into the next tile size sooner), so wider keys keep 32.
    sh.m = sh.parts * (sh.p2 > 1 ? (esz == 4 ? 16 : 32) : 64);
    return sh;
}

}  // namespace

// Sorts data[0, n) into out (out may equal data) with `values` as scratch of
// at least n elements and the call's other temporaries from `tmp`.  Sample
// sort into buckets of ~3/8 of the on-chip tile: one partition level
// (data -> values) for up to kPartOneLevel parts, else two (data -> values ->
// out) with p1 x p2 parts taken from one global splitter list, so the second
// level needs no sampling of its own.  The buckets are then sorted as one
// ragged batch (values -> out, or in place in out), planned on the device;
// the rare bucket longer than a tile sends the call to a host-planned batch
// that cuts it into chunks merged by tiled passes.  Host synchronisations:
// first-level bucket starts (two levels only) and the final plan's header;
// the host-planned path reads all bucket starts and synchronises once more at
// its end.
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

    unsigned char *samples = nullptr, *samples_tmp = nullptr, *gsp = nullptr, *l1sp = nullptr;
    int rc = tmp.get(&samples, size_t(m) * esz);
    if (rc == CS_OK) rc = tmp.get(&samples_tmp, size_t(m) * esz);
    if (rc == CS_OK) rc = tmp.get(&gsp, size_t(parts - 1) * size_t(pk.key_size));
    if (rc == CS_OK && p2 > 1) rc = tmp.get(&l1sp, size_t(std::max(p1 - 1, 1)) * size_t(pk.key_size));
    if (rc != CS_OK) return rc;

    // splitters from a sorted hashed sample; the first level's are every
    // p2-th of them
    hipError_t e = pk.sample(data, n, samples, int(m), stream);
    if (e != hipSuccess) return int(e);
    rc = merge_sort(dtype, samples, m, samples, samples_tmp, tmp);
    if (rc != CS_OK) return rc;
    e = pk.splitters(samples, int(m), int(parts), gsp, stream);
    if (e == hipSuccess && p2 > 1) e = pk.gather(gsp, p1, p2, l1sp, stream);
    if (e != hipSuccess) return int(e);

    // level 1: data -> values
    thread_local PartHost host;
    Level& L1 = host.l1;
    Level& L2 = host.l2;
    clear_level(L1);
    clear_level(L2);
    add_segment(L1, 0, n, 0, p1 - 1, -1, -1, false);
    std::vector<int64_t>& st1 = host.st1;
    int64_t *d_st1 = nullptr, *d_st2 = nullptr;
    unsigned char *d_eq1 = nullptr, *d_eq2 = nullptr;
    rc = tmp.get(&d_st1, size_t(L1.nbuckets) + 1);
    if (rc == CS_OK) rc = tmp.get(&d_eq1, size_t(L1.nbuckets));
    if (rc == CS_OK)
        rc = run_level(pk, L1, data, p2 > 1 ? l1sp : gsp, values, d_st1, d_eq1, p2 > 1 ? &st1 : nullptr, tmp);
    if (rc != CS_OK) return rc;

    // level 2: values -> out, one segment per first-level bucket; equality
    // buckets and buckets already small are copied through
    const Level* fin = &L1;
    const std::vector<int64_t>* fst = &st1;
    std::vector<int64_t>& st2 = host.st2;
    void* src = values;
    if (p2 > 1) {
        const int nb1 = 2 * p1 - 1;
        for (int b = 0; b < nb1; ++b) {
            const int64_t b0 = st1[size_t(b)], b1 = b + 1 < nb1 ? st1[size_t(b) + 1] : n;
            const int j = b / 2;
            if ((b & 1) || b1 - b0 <= target) {
                add_segment(L2, b0, b1, 0, 0, -1, -1, (b & 1) != 0);
            } else {
                add_segment(L2, b0, b1, j * p2, p2 - 1, j > 0 ? j * p2 - 1 : -1,
                            j < p1 - 1 ? (j + 1) * p2 - 1 : -1, false);
            }
        }
        rc = tmp.get(&d_st2, size_t(L2.nbuckets) + 1);
        if (rc == CS_OK) rc = tmp.get(&d_eq2, size_t(L2.nbuckets));
        if (rc == CS_OK) rc = run_level(pk, L2, values, gsp, out, d_st2, d_eq2, nullptr, tmp);
        if (rc != CS_OK) return rc;
        fin = &L2;
        fst = &st2;
        src = out;
    }
    int64_t* d_fst = p2 > 1 ? d_st2 : d_st1;
    const unsigned char* d_feq = p2 > 1 ? d_eq2 : d_eq1;
    const int64_t fnb = fin->nbuckets;

    // plan the final ragged sort on the device; only its header comes back
    {
        const SegClasses sc = seg_classes(dtype);
        PlanTable t{};
        t.maxb = sc.maxb;
        t.small_max = sc.small_max;
        for (int c = 0; c < kPlanClasses; ++c) {
            t.spi[c] = c <= sc.maxb ? std::max(1, sc.spi[c]) : 1;
            t.variant[c] = c <= sc.maxb ? sc.variant[c] : 0;
        }
        unsigned char* cls = nullptr;
        int *meta = nullptr, *d_list = nullptr, *d_counters = nullptr;
        SegItem* d_items = nullptr;
        rc = tmp.get(&cls, size_t(fnb));
        if (rc == CS_OK) rc = tmp.get(&meta, kPlanMeta);
        if (rc == CS_OK) rc = tmp.get(&d_list, size_t(fnb));
        if (rc == CS_OK) rc = tmp.get(&d_items, size_t(fnb));
        if (rc == CS_OK) rc = tmp.get(&d_counters, 2);
        if (rc != CS_OK) return rc;
        const unsigned g = unsigned((fnb + 255) / 256);
        hipLaunchKernelGGL(k_plan_zero, dim3(1), dim3(128), 0, stream, meta);
        // equality buckets need no sort, but an out-of-place final sort must still move them
        const int skip_equal = src == out ? 1 : 0;
        hipLaunchKernelGGL(k_plan_classify, dim3(g), dim3(256), 0, stream, d_fst, d_feq, skip_equal, fnb, n, lk.chunk,
                           cls, meta);
        hipLaunchKernelGGL(k_plan_layout, dim3(1), dim3(64), 0, stream, meta, t);
        hipLaunchKernelGGL(k_plan_fill, dim3(g), dim3(256), 0, stream, cls, fnb, meta, d_list);
        hipLaunchKernelGGL(k_plan_items, dim3(g), dim3(256), 0, stream, meta, t, d_items);
        int hdr[3] = {0, 0, 0};
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(hdr, meta + 4 * kPlanClasses, sizeof(hdr), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return int(e);
        if (hdr[2] == 0) {   // no bucket longer than the tile: sort with the device plan
            SegShape ss;
            ss.grid[kSegMulti] = seg_persistent_grid(dtype, kSegMulti, tmp.grid_cap);
            ss.grid[kSegSingle] = seg_persistent_grid(dtype, kSegSingle, tmp.grid_cap);
            if (ss.grid[kSegMulti] <= 0 || ss.grid[kSegSingle] <= 0) return CS_E_NO_DEVICE;
            e = seg_launch_all(dtype, ss, src, out, d_fst, d_list, d_items, 0, d_counters, stream,
                               meta + 4 * kPlanClasses);
            return e == hipSuccess ? CS_OK : int(e);
        }
    }
    // some bucket is longer than the tile: the host plans the sort and
    // cuts the long buckets into chunks
    std::vector<int64_t>& fst_host = p2 > 1 ? st2 : st1;
    fst_host.assign(size_t(fnb), 0);
    e = hipMemcpyAsync(fst_host.data(), d_fst, size_t(fnb) * sizeof(int64_t), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return int(e);

    // one ragged batch over all buckets (src -> out); long buckets cut into
    // chunks.  The level's bucket starts run through the array in order.
    const int64_t chunk = lk.chunk;
    const std::vector<int64_t>& st = *fst;
    std::vector<int64_t>& seg = host.seg;
    seg.assign(1, 0);
    seg.reserve(st.size() / 2 + size_t(n / chunk) + 2);
    std::vector<MergeTile>& tiles = host.tiles;
    tiles.clear();
    int passes = 0;
    size_t g = 0;   // segment of bucket i (only needed for long buckets)
    for (size_t i = 0; i < st.size(); ++i) {
        const int64_t b0 = st[i], b1 = i + 1 < st.size() ? st[i + 1] : n;
        const int64_t len = b1 - b0;
        if (len <= 0) continue;
        if (b0 != seg.back()) return CS_E_NO_DEVICE;   // buckets must tile the array in order
        if (len <= chunk) {
            seg.push_back(b1);
            continue;
        }
        for (int64_t c = 0; c < len; c += chunk) seg.push_back(b0 + std::min(len, c + chunk));
        while (g + 1 < fin->segs.size() && fin->segs[g + 1].first <= int64_t(i)) ++g;
        const bool equal = fin->seg_equal[g] || ((int64_t(i) - fin->segs[g].first) & 1);
        if (!equal) {   // equality buckets are sorted once their chunks are
            int p = 0;
            for (int64_t run = chunk; run < len; run *= 2) ++p;
            passes = std::max(passes, p);
            for (int64_t d0 = 0; d0 < len; d0 += lk.tile) tiles.push_back(MergeTile{b0, len, d0});
        }
    }
    if (seg.back() != n) return CS_E_NO_DEVICE;
    std::vector<int>& list = host.list;
    std::vector<SegItem>& items = host.items;
    const int64_t nseg = int64_t(seg.size()) - 1;
    rc = seg_build_items(dtype, seg.data(), nseg, nullptr, list, items, nullptr, nullptr, nullptr);
    if (rc != CS_OK) return rc;
    const SegShape shape = seg_shape(dtype, items);
    int64_t* d_seg = nullptr;
    int *d_list = nullptr, *d_counters = nullptr;
    SegItem* d_items = nullptr;
    MergeTile* d_tiles = nullptr;
    rc = tmp.get(&d_seg, seg.size());
    if (rc == CS_OK) rc = tmp.get(&d_list, list.size());
    if (rc == CS_OK) rc = tmp.get(&d_items, items.size());
    if (rc == CS_OK) rc = tmp.get(&d_counters, 2);
    if (rc == CS_OK && !tiles.empty()) rc = tmp.get(&d_tiles, tiles.size());
    if (rc != CS_OK) return rc;
    e = hipMemcpyAsync(d_seg, seg.data(), seg.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_list, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(SegItem), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(MergeTile), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = seg_launch_all(dtype, shape, src, out, d_seg, d_list, d_items, int(items.size()), d_counters, stream);
    rc = e == hipSuccess ? CS_OK : int(e);
    // merge passes over the long buckets (out <-> values), then final placement
    void* a = out;
    void* b2 = values;
    for (int p = 0; rc == CS_OK && p < passes; ++p) {
        e = pk.seg_merge(a, b2, d_tiles, int(tiles.size()), chunk << p, 0, stream);
        if (e != hipSuccess) rc = int(e);
        std::swap(a, b2);
    }
    if (rc == CS_OK && a != out) {
        e = pk.seg_merge(a, out, d_tiles, int(tiles.size()), 0, 1, stream);
        if (e != hipSuccess) rc = int(e);
    }
    const hipError_t se = hipStreamSynchronize(stream);   // host arrays above are read by the copies
    if (rc == CS_OK && se != hipSuccess) rc = int(se);
    return rc;
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
    total += 2 * r(size_t(sh.m) * esz);                                   // samples, sorted copy
    total += merge_scratch_bytes(dtype, sh.m);                            // their merge sort
    total += r(size_t(sh.parts - 1) * size_t(pk.key_size));               // global splitters
    total += r(size_t(std::max(sh.p1 - 1, 1)) * size_t(pk.key_size));     // first-level splitters
    auto level = [&](int64_t segs, int64_t blocks, int64_t nbuckets, int64_t cells) {
        size_t b = r(size_t(nbuckets + 1) * 8) + r(size_t(nbuckets));     // starts, equality flags
        b += r(size_t(blocks) * sizeof(PartBlock)) + r(size_t(segs) * sizeof(PartSeg));
        b += r(size_t(cells) * 8) + r(size_t((cells + kScanTile - 1) / kScanTile + 1) * 8);
        return b;
    };
    const int64_t blocks1 = (n + kPartBlockElems - 1) / kPartBlockElems;
    const int64_t nb1 = 2 * int64_t(sh.p1) - 1;
    total += level(1, blocks1, nb1, nb1 * blocks1);
    int64_t fnb = nb1;
    if (sh.p2 > 1) {
        const int64_t nbs = 2 * int64_t(sh.p2) - 1;                       // buckets per second-level segment
        const int64_t blocks2 = blocks1 + nb1;                            // each segment rounds up once
        fnb = nbs * nb1;
        total += level(nb1, blocks2, fnb, nbs * blocks2);
    }
    // device plan of the final sort
    total += r(size_t(fnb)) + r(size_t(kPlanMeta) * 4) + r(size_t(fnb) * 4) + r(size_t(fnb) * sizeof(SegItem)) + r(8);
    // host-planned fallback: every bucket plus every chunk of a long one
    const int64_t segs = fnb + n / lk.chunk + 2;
    total += r(size_t(segs) * 8) + r(size_t(segs) * 4) + r(size_t(segs) * sizeof(SegItem)) + r(8);
    total += r(size_t(n / lk.tile + fnb + 1) * sizeof(MergeTile));
    return total;
}

}  // namespace cs
