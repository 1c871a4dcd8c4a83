// Ragged batches: many arrays of uneven length given as CSR offsets
// (segment s = [off[s], off[s+1])).
//
// Persistent bucketed kernel.  At plan time every segment is put in the
// bucket b = ceil(log2(len)) (W = 2^b) and the segments of a bucket are cut
// into work items of as many segments as one workgroup of 256 threads sorts
// at once with that bucket's tiling.  Items are listed largest bucket
// first.  Two persistent kernels share the list: one for the buckets whose
// register tiles span several waves (and their LDS-resident variant), one
// for the buckets whose tiles fit one wave.  Each is compiled for its own
// buckets only, so its register and LDS footprint (and with it the number
// of workgroups per CU) is set by those buckets and not by the largest
// tile.  Each launch (a grid of a few workgroups per CU) pulls items of its
// range from its own global counter until the range is exhausted; each
// item runs the tile sort of its bucket with the variant the tuner chose:
//   variant 0: register tile, 32 keys per lane (16 for 16-byte records)
//   variant 1: register tile, 16 keys per lane
//   variant 2: LDS-resident network, one segment per workgroup
//   variant 3: cross-lane register engine (bitonic_xlane.hpp), 4-byte keys
//              without payload, W = 32..2048
// Every wave reaches the loop exit once the counter passes the item count.
//
// The reference's only ragged sort is SortPartitions (cusort.hpp:1802-1846):
// one block per bucket with LDS sized for the largest bucket.
// k_seg_plain is that baseline on this hardware: one lane group per
// segment, every segment padded to the maximum length's tiling.
#pragma once

#include "batched.hpp"

namespace cs {

constexpr int kSegBlock = 256;      // threads per workgroup of the segmented kernels
constexpr int kSegVariants = 4;

struct SegItem {
    int bucket;    // log2 W
    int variant;   // 0 .. 3 (see above)
    int first;     // index into the bucket-ordered segment list
    int count;     // segments in this item
};

template <int DT> constexpr int seg_max_logw() {
    return sizeof(typename Elem<DT>::Storage) == 4 ? 13 : (sizeof(typename Elem<DT>::Storage) == 8 ? 12 : 12);
}

// Tiling (log2 E) of a register-tile variant for bucket b; -1 if absent.
// Variant 0 is the batched sort's default tiling.
template <int DT> constexpr int seg_loge(int b, int variant) {
    if (variant == 0) return auto_loge(DT, b);
    if (variant == 1) return b < 4 ? -1 : 4;     // distinct from variant 0 only when b >= 5
    return -1;
}

template <int DT> constexpr bool seg_variant_ok(int b, int variant) {
    if (b < 0 || b > seg_max_logw<DT>()) return false;
    if (variant == 2) return b >= 1;
    if (variant == 3) return xlane_dtype(DT) && b >= xl::kLogE && b <= 11;
    const int le = seg_loge<DT>(b, variant);
    if (le < 0 || (variant == 1 && b < 5)) return false;
    if (b > 0 && le == 0) return false;
    return (1 << (b - le)) <= kSegBlock;
}

template <int DT> constexpr int seg_segments_per_item(int b, int variant) {
    if (variant == 2) return 1;
    if (variant == 3) return kSegBlock >> (b - xl::kLogE);
    const int T = 1 << (b - seg_loge<DT>(b, variant));
    return kSegBlock / T;
}

template <class E> constexpr int tile_lds_words(int logw, int loge) {
    const int T = 1 << (logw - loge);
    const int apb = (T >= 256 ? T : 256) / T;
    const int padshift = sizeof(typename E::Key) == 4 ? 5 : 4;
    const int region = (1 << logw) + ((1 << logw) >> padshift);
    return make_plan(logw, loge).ntrans ? (T > 64 ? 2 : 1) * apb * region : 0;
}

// Largest bucket of the single-wave kernel: variant 0 tiles of at most 64
// lanes.  The multi-wave kernel takes the buckets above it.
template <int DT> constexpr int seg_small_max() {
    int m = 0;
    for (int b = 0; b < seg_max_logw<DT>(); ++b)
        if (b - seg_loge<DT>(b, 0) <= 6) m = b;
    return m;
}

// LDS words (keys; payload uses the same count of int32) a persistent
// kernel for buckets BLO..BHI needs: the maximum over their variants.
template <int DT, int BLO, int BHI> constexpr int seg_lds_words() {
    int m = 1 << BHI;    // LDS-resident variant: one segment of W
    for (int b = BLO; b <= BHI; ++b) {
        for (int v = 0; v < 2; ++v)
            if (seg_variant_ok<DT>(b, v)) {
                const int w = tile_lds_words<Elem<DT>>(b, seg_loge<DT>(b, v));
                m = w > m ? w : m;
            }
        if (seg_variant_ok<DT>(b, 3)) {
            const int w = seg_segments_per_item<DT>(b, 3) * xl::region_words(b);
            m = w > m ? w : m;
        }
    }
    return m;
}

// LDS-resident network for one segment with all threads of the block.
template <class E>
__device__ __forceinline__ void lds_segment_sort(const typename E::Storage* src, typename E::Storage* dst, int n, int logw,
                                 typename E::Key* lk, int32_t* lp) {
    const int W = 1 << logw;
    const int T = int(blockDim.x);
    const int tid = int(threadIdx.x);
    for (int i = tid; i < W; i += T) {
        if (i < n) {
            const typename E::Storage s = src[i];
            lk[i] = E::key(s);
            if constexpr (E::kPayload) lp[i] = E::payload(s);
        } else {
            lk[i] = E::kPad;
            if constexpr (E::kPayload) lp[i] = INT32_MAX;
        }
    }
    __syncthreads();
    for (int k = 1; k <= logw; ++k) {
        const int half = 1 << (k - 1);
        for (int q = tid; q < W / 2; q += T) {
            const int i = ((q >> (k - 1)) << k) | (q & (half - 1));
            lds_cx<E>(lk, lp, i, i ^ ((half << 1) - 1));
        }
        __syncthreads();
        for (int d = half >> 1; d >= 1; d >>= 1) {
            const int ld = __builtin_ctz(d);
            for (int q = tid; q < W / 2; q += T) {
                const int i = ((q >> ld) << (ld + 1)) | (q & (d - 1));
                lds_cx<E>(lk, lp, i, i + d);
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < n; i += T) dst[i] = E::make(lk[i], E::kPayload ? lp[i] : 0);
    __syncthreads();
}

// A list entry names a segment s (off[s] .. off[s+1]) or one half of it: a
// planner may cut a segment a little longer than 2^small_max into its head
// (the first 2^small_max elements, kSegHead) and its tail (the rest,
// kSegTail), so that both halves take the single-wave tiles; the two sorted
// halves are merged afterwards.
constexpr int kSegHead = 1 << 30;
constexpr int kSegTail = 1 << 29;
constexpr int kSegIndex = kSegTail - 1;

template <int DT>
__device__ __forceinline__ void seg_range(const int64_t* off, int entry, int64_t& b0, int& len) {
    constexpr int64_t H = int64_t(1) << seg_small_max<DT>();
    const int s = entry & kSegIndex;
    b0 = off[s];
    int64_t l = off[s + 1] - b0;
    if (entry & kSegHead) l = H;
    if (entry & kSegTail) {
        b0 += H;
        l -= H;
    }
    len = int(l);
}

// The body of every (bucket, variant) is a separate, non-inlined function so
// that each gets its own register allocation: inlined into one kernel, the
// allocator sized the whole persistent kernel for its worst bucket (more
// than 400 registers per lane, one wave per SIMD).  Their global pointers
// carry the global address space explicitly, so that the loads and stores
// inside stay global (not flat) instructions.
template <class T> using global_ptr = __attribute__((address_space(1))) T*;
template <class T> __device__ __forceinline__ global_ptr<T> to_global(T* p) { return (global_ptr<T>)p; }

// One item: `count` segments of bucket B (list entries first..first+count-1)
// with register-tile variant V.
template <int DT, int B, int VAR>
__device__ __attribute__((noinline)) void seg_tile_item(global_ptr<const typename Elem<DT>::Storage> gin,
                                                        global_ptr<typename Elem<DT>::Storage> gout,
                                                        global_ptr<const int64_t> goff, global_ptr<const int> glist,
                                                        int first, int count, typename Elem<DT>::Key* lk,
                                                        int32_t* lp) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    const S* in = (const S*)gin;
    S* out = (S*)gout;
    const int64_t* off = (const int64_t*)goff;
    const int* list = (const int*)glist;
    constexpr int LE = seg_loge<DT>(B, VAR);
    using C = TileCfg<E, B, LE>;
    const int tid = int(threadIdx.x);
    const int group = tid >> C::LOGT;
    const int t = tid & (C::T - 1);
    const bool active = group < count;
    int64_t b0 = 0;
    int len = 0;
    if (active) seg_range<DT>(off, list[first + group], b0, len);
    tile_sort<E, B, LE>(in + b0, out + b0, len, t, lk + (C::NTRANS ? group * C::REGION : 0),
                        lp + ((C::NTRANS && E::kPayload) ? group * C::REGION : 0));
}

// One item of the cross-lane variant: `count` segments of bucket B, segment
// i of the item on the lanes of array slot i (lane mapping of the batched
// cross-lane kernel), each slot with its own LDS region.
template <int DT, int B>
__device__ __attribute__((noinline)) void seg_xlane_item(global_ptr<const typename Elem<DT>::Storage> gin,
                                                         global_ptr<typename Elem<DT>::Storage> gout,
                                                         global_ptr<const int64_t> goff, global_ptr<const int> glist,
                                                         int first, int count, typename Elem<DT>::Key* lk) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    constexpr int LOGT = B - xl::kLogE;
    constexpr int APW = 64 >> LOGT;
    constexpr int REGION = xl::region_words(B);
    const int tid = int(threadIdx.x);
    const int p = tid & 63;
    const int t = xl::lane_t(p, LOGT);
    const int slot = (tid >> 6) * APW + xl::lane_g(p, LOGT);
    int64_t b0 = 0;
    int len = 0;
    if (slot < count) seg_range<DT>((const int64_t*)goff, ((const int*)glist)[first + slot], b0, len);
    xlane_sort<E, B, kXlaneWindow, true>((const S*)gin + b0, (S*)gout + b0, len, t, lk + slot * REGION);
}

// One item of the LDS-resident variant: the single segment list[first].
template <int DT, int B>
__device__ __attribute__((noinline)) void seg_lds_item(global_ptr<const typename Elem<DT>::Storage> gin,
                                                       global_ptr<typename Elem<DT>::Storage> gout,
                                                       global_ptr<const int64_t> goff, global_ptr<const int> glist,
                                                       int first, typename Elem<DT>::Key* lk, int32_t* lp) {
    using S = typename Elem<DT>::Storage;
    int64_t b0 = 0;
    int len = 0;
    seg_range<DT>((const int64_t*)goff, ((const int*)glist)[first], b0, len);
    lds_segment_sort<Elem<DT>>((const S*)gin + b0, (S*)gout + b0, len, B, lk, lp);
}

template <int DT, int B>
__device__ __forceinline__ void seg_bucket(int variant, const typename Elem<DT>::Storage* in,
                                           typename Elem<DT>::Storage* out, const int64_t* off, const int* list,
                                           int first, int count, typename Elem<DT>::Key* lk, int32_t* lp) {
    if constexpr (seg_variant_ok<DT>(B, 0)) {
        if (variant == 0) {
            seg_tile_item<DT, B, 0>(to_global(in), to_global(out), to_global(off), to_global(list), first, count,
                                    lk, lp);
            return;
        }
    }
    if constexpr (seg_variant_ok<DT>(B, 1)) {
        if (variant == 1) {
            seg_tile_item<DT, B, 1>(to_global(in), to_global(out), to_global(off), to_global(list), first, count,
                                    lk, lp);
            return;
        }
    }
    if constexpr (seg_variant_ok<DT>(B, 3)) {
        if (variant == 3) {
            seg_xlane_item<DT, B>(to_global(in), to_global(out), to_global(off), to_global(list), first, count, lk);
            return;
        }
    }
    if constexpr (seg_variant_ok<DT>(B, 2)) {
        if (variant == 2)
            seg_lds_item<DT, B>(to_global(in), to_global(out), to_global(off), to_global(list), first, lk, lp);
    }
}

template <int DT, int B, int BHI>
__device__ __forceinline__ void seg_dispatch(int bucket, int variant, const typename Elem<DT>::Storage* in,
                                             typename Elem<DT>::Storage* out, const int64_t* off, const int* list,
                                             int first, int count, typename Elem<DT>::Key* lk, int32_t* lp) {
    if constexpr (B <= BHI) {
        if (bucket == B) {
            seg_bucket<DT, B>(variant, in, out, off, list, first, count, lk, lp);
            return;
        }
        seg_dispatch<DT, B + 1, BHI>(bucket, variant, in, out, off, list, first, count, lk, lp);
    }
}

// Workgroups of a persistent kernel for buckets BLO..BHI that fit a CU's
// 160 KB of LDS; each has one wave per SIMD, so this is also the number of
// waves per SIMD the register allocation must leave room for (left alone,
// the allocator sizes the kernel for one wave per SIMD).
template <int DT, int BLO, int BHI> constexpr int seg_waves_per_simd() {
    const int bytes = seg_lds_words<DT, BLO, BHI>() * int(sizeof(typename Elem<DT>::Key) +
                                                          (Elem<DT>::kPayload ? 4 : 0));
    const int w = (160 * 1024) / (bytes + 64);
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// Persistent kernel for the items [begin, end) of the list, whose buckets
// all lie in BLO..BHI; *counter starts at 0.  With `range` (a plan built on
// the device) the item range is read from it instead: range[0] items of the
// multi-wave kernel (BLO > 0) come first, then the single-wave kernel's up
// to range[1].
template <int DT, int BLO, int BHI>
__global__ void __launch_bounds__(kSegBlock, (seg_waves_per_simd<DT, BLO, BHI>()))
k_seg_persistent(const typename Elem<DT>::Storage* in, typename Elem<DT>::Storage* out, const int64_t* off,
                 const int* list, const SegItem* items, int begin, int end, int* counter, const int* range) {
    using E = Elem<DT>;
    if (range) {
        begin = BLO > 0 ? 0 : range[0];
        end = BLO > 0 ? range[0] : range[1];
    }
    __shared__ typename E::Key lds_k[seg_lds_words<DT, BLO, BHI>()];
    __shared__ int32_t lds_p[E::kPayload ? seg_lds_words<DT, BLO, BHI>() : 1];
    __shared__ int item_slot;
    for (;;) {
        if (threadIdx.x == 0) item_slot = begin + atomicAdd(counter, 1);
        __syncthreads();
        const int i = item_slot;
        __syncthreads();                       // everyone has read the slot before it is reused
        if (i >= end) break;                   // uniform: every wave leaves together
        const SegItem it = items[i];
        seg_dispatch<DT, BLO, BHI>(it.bucket, it.variant, in, out, off, list, it.first, it.count, lds_k, lds_p);
    }
}

// Baseline: one lane group per segment, every segment in the tiling of the
// maximum bucket (the reference's SortPartitions pattern).
template <int DT, int B>
__global__ void __launch_bounds__(kSegBlock)
k_seg_plain(const typename Elem<DT>::Storage* in, typename Elem<DT>::Storage* out, const int64_t* off,
            int64_t segments) {
    using E = Elem<DT>;
    constexpr int LE = seg_loge<DT>(B, 0);
    using C = TileCfg<E, B, LE>;
    __shared__ typename E::Key lds_k[C::LDS_KEYS];
    __shared__ int32_t lds_p[C::LDS_PAYLOAD];
    const int tid = int(threadIdx.x);
    const int group = tid >> C::LOGT;
    const int t = tid & (C::T - 1);
    const int64_t s = int64_t(blockIdx.x) * C::APB + group;
    const bool active = s < segments;
    const int64_t b0 = active ? off[s] : 0;
    const int len = active ? int(off[s + 1] - b0) : 0;
    tile_sort<E, B, LE>(in + b0, out + b0, len, t, lds_k + (C::NTRANS ? group * C::REGION : 0),
                        lds_p + ((C::NTRANS && E::kPayload) ? group * C::REGION : 0));
}

// Zeroes the two item counters of the persistent kernels.
__global__ void k_seg_reset(int* counters);

// The two persistent kernels: kSegMulti for buckets above small_max (tiles
// of several waves), kSegSingle for buckets 0..small_max.
constexpr int kSegMulti = 0;
constexpr int kSegSingle = 1;

// Host side (segmented_<dtype>.hip).
struct SegKernels {
    hipError_t (*persistent)(int which, const void* in, void* out, const int64_t* off, const int* list,
                             const SegItem* items, int begin, int end, int* counter, const int* range, int grid,
                             hipStream_t s) = nullptr;
    const void* persistent_fn[2] = {nullptr, nullptr};
    int small_max = 0;   // largest bucket of the kSegSingle kernel
    hipError_t (*plain)(const void* in, void* out, const int64_t* off, int64_t segments, int bucket,
                        hipStream_t s) = nullptr;
    int max_logw = 0;
    int (*segments_per_item)(int bucket, int variant) = nullptr;
    bool (*variant_ok)(int bucket, int variant) = nullptr;
};

template <int DT>
hipError_t seg_launch_persistent(int which, const void* in, void* out, const int64_t* off, const int* list,
                                 const SegItem* items, int begin, int end, int* counter, const int* range, int grid,
                                 hipStream_t s) {
    using S = typename Elem<DT>::Storage;
    constexpr int SM = seg_small_max<DT>();
    if ((!range && end <= begin) || grid <= 0) return hipSuccess;
    if (which == kSegMulti)
        hipLaunchKernelGGL((k_seg_persistent<DT, SM + 1, seg_max_logw<DT>()>), dim3(unsigned(grid)), dim3(kSegBlock),
                           0, s, static_cast<const S*>(in), static_cast<S*>(out), off, list, items, begin, end,
                           counter, range);
    else
        hipLaunchKernelGGL((k_seg_persistent<DT, 0, SM>), dim3(unsigned(grid)), dim3(kSegBlock), 0, s,
                           static_cast<const S*>(in), static_cast<S*>(out), off, list, items, begin, end, counter,
                           range);
    return hipGetLastError();
}

template <int DT>
hipError_t seg_launch_plain(const void* in, void* out, const int64_t* off, int64_t segments, int bucket,
                            hipStream_t s) {
    using S = typename Elem<DT>::Storage;
    hipError_t err = hipErrorInvalidValue;
    hfor<0, seg_max_logw<DT>() + 1>([&](auto b_) {
        constexpr int B = decltype(b_)::value;
        if (B == bucket) {
            using C = TileCfg<Elem<DT>, B, seg_loge<DT>(B, 0)>;
            const int64_t blocks = (segments + C::APB - 1) / C::APB;
            if (blocks > 0)
                hipLaunchKernelGGL((k_seg_plain<DT, B>), dim3(unsigned(blocks)), dim3(kSegBlock), 0, s,
                                   static_cast<const S*>(in), static_cast<S*>(out), off, segments);
            err = hipGetLastError();
        }
    });
    return err;
}

template <int DT> int seg_spi(int b, int v) { return seg_segments_per_item<DT>(b, v); }
template <int DT> bool seg_ok(int b, int v) { return seg_variant_ok<DT>(b, v); }

template <int DT>
SegKernels seg_kernels() {
    SegKernels k;
    constexpr int SM = seg_small_max<DT>();
    k.persistent = &seg_launch_persistent<DT>;
    k.persistent_fn[kSegMulti] =
        reinterpret_cast<const void*>(&k_seg_persistent<DT, SM + 1, seg_max_logw<DT>()>);
    k.persistent_fn[kSegSingle] = reinterpret_cast<const void*>(&k_seg_persistent<DT, 0, SM>);
    k.small_max = SM;
    k.plain = &seg_launch_plain<DT>;
    k.max_logw = seg_max_logw<DT>();
    k.segments_per_item = &seg_spi<DT>;
    k.variant_ok = &seg_ok<DT>;
    return k;
}

SegKernels seg_kernels_for(int dtype);

}  // namespace cs
