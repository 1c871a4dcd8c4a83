// Large-array sort: chunk sort + merge-path merge passes (the reference's
// SortMerge, cusort.hpp:2025-2300), and the sample-sort partition kernels
// (the reference's Sort / HybridSort partitioned path, cusort.hpp:1303-2023,
// 2488-2754), redesigned for gfx950.
//
// Merge pass.  Runs of length `run` are merged pairwise into runs of 2*run.
// Each workgroup (256 threads) produces one output tile of TILE = 256*IPT
// elements.  Its two co-ranks (how many elements of the left run precede
// the tile's first and one-past-last output) are found by one wave each with
// a 64-ary search: every lane probes one point of the remaining interval in
// global memory and a ballot keeps the sub-interval that holds the answer,
// so a run of 2^27 needs five probe rounds instead of 27 dependent loads.
// The two input slices are staged in LDS with coalesced loads, every thread
// finds its own co-rank inside LDS and merges IPT outputs sequentially, and
// the tile leaves through LDS as lane-striped 16-byte stores.
//
// Ties take the left run first, so the merge is stable (the reference's
// merge takes A on ties as well, cusort.hpp:2120).
#pragma once

#include "batched.hpp"

namespace cs {

constexpr int kMergeThreads = 256;

template <int DT> constexpr int merge_ipt() { return sizeof(typename Elem<DT>::Storage) == 4 ? 16 : 8; }
template <int DT> constexpr int merge_tile() { return kMergeThreads * merge_ipt<DT>(); }
// Chunk length of the initial tile sort (one workgroup per chunk).
template <int DT> constexpr int chunk_logw() { return sizeof(typename Elem<DT>::Storage) == 4 ? 13 : 12; }

// Number of elements of A among the first d outputs of merge(A, B), ties
// taking A first: the smallest i with !(A[i] <= B[d-1-i]) in the feasible
// range [max(0, d-nb), min(d, na)].  Whole wave cooperates; every lane
// returns the same value.
template <class E>
__device__ int64_t wave_corank(const typename E::Storage* A, int64_t na, const typename E::Storage* B, int64_t nb,
                               int64_t d) {
    const int lane = int(threadIdx.x & 63);
    int64_t lo = d > nb ? d - nb : 0;
    int64_t hi = d < na ? d : na;
    while (hi > lo) {
        const int64_t span = hi - lo;
        const int probes = span < 64 ? int(span) : 64;
        const int64_t m = lo + (int64_t(lane) * span) / probes;   // probes distinct, all < hi
        bool p = false;
        if (lane < probes) p = !(E::key(B[d - 1 - m]) < E::key(A[m]));   // A[m] <= B[d-1-m]
        const uint64_t ball = __ballot(p);
        const int c = __popcll(ball);       // predicate holds on a prefix of the probes
        if (c == 0) {
            hi = lo;
        } else {
            const int64_t mc1 = lo + (int64_t(c - 1) * span) / probes;
            const int64_t nh = c < probes ? lo + (int64_t(c) * span) / probes : hi;
            lo = mc1 + 1;
            hi = nh;
        }
    }
    return lo;
}

// Same search inside LDS for one thread (binary).
template <class E>
__device__ __forceinline__ int lds_corank(const typename E::Storage* A, int na, const typename E::Storage* B, int nb,
                                          int d) {
    int lo = d > nb ? d - nb : 0;
    int hi = d < na ? d : na;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (!(E::key(B[d - 1 - mid]) < E::key(A[mid]))) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One output tile [g0, g0 + TILE) of a merge pass over the array src[0, n)
// whose runs have length `run`; all kMergeThreads threads take part.
template <int DT>
__device__ __forceinline__ void merge_tile_body(const typename Elem<DT>::Storage* src,
                                                typename Elem<DT>::Storage* dst, int64_t n, int64_t run,
                                                int64_t g0) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    constexpr int IPT = merge_ipt<DT>();
    constexpr int TILE = merge_tile<DT>();
    constexpr int V = 16 / int(sizeof(S)) > 0 ? 16 / int(sizeof(S)) : 1;
    constexpr int PAD = sizeof(S) == 4 ? 5 : (sizeof(S) == 8 ? 4 : 3);
    __shared__ S lds_in[TILE];
    __shared__ S lds_out[TILE + (TILE >> PAD)];
    __shared__ int64_t corank[2];

    if (g0 >= n) return;
    const int64_t base = (g0 / (2 * run)) * (2 * run);
    const int64_t na = run < n - base ? run : n - base;
    const int64_t rem = n - base - na;
    const int64_t nb = rem < 0 ? 0 : (rem < run ? rem : run);
    const S* A = src + base;
    const S* B = A + na;
    const int64_t d0 = g0 - base;
    const int64_t d1 = d0 + TILE < na + nb ? d0 + TILE : na + nb;
    const int tid = int(threadIdx.x);
    const int wave = tid >> 6;
    if (wave < 2) {
        const int64_t c = wave_corank<E>(A, na, B, nb, wave == 0 ? d0 : d1);
        if ((tid & 63) == 0) corank[wave] = c;
    }
    __syncthreads();
    const int64_t a0 = corank[0], a1 = corank[1];
    const int la = int(a1 - a0);
    const int lb = int((d1 - a1) - (d0 - a0));
    const int64_t b0 = d0 - a0;
    for (int i = tid; i < la; i += kMergeThreads) lds_in[i] = A[a0 + i];
    for (int i = tid; i < lb; i += kMergeThreads) lds_in[la + i] = B[b0 + i];
    __syncthreads();
    const S* sA = lds_in;
    const S* sB = lds_in + la;
    const int total = la + lb;
    const int dt0 = tid * IPT < total ? tid * IPT : total;
    int i = lds_corank<E>(sA, la, sB, lb, dt0);
    int j = dt0 - i;
    S cur_a = i < la ? sA[i] : S{};
    S cur_b = j < lb ? sB[j] : S{};
    #pragma unroll
    for (int k = 0; k < IPT; ++k) {
        if (dt0 + k < total) {
            const bool take_a = j >= lb || (i < la && !(E::key(cur_b) < E::key(cur_a)));
            const int x = dt0 + k;
            lds_out[x + (x >> PAD)] = take_a ? cur_a : cur_b;
            if (take_a) { ++i; if (i < la) cur_a = sA[i]; }
            else { ++j; if (j < lb) cur_b = sB[j]; }
        }
    }
    __syncthreads();
    S* out = dst + base + d0;
    const bool vec = (((uintptr_t)out) & 15) == 0 && V > 1;
    for (int q = tid * V; q < total; q += kMergeThreads * V) {
        if (vec && q + V <= total) {
            S tmp[V];
            #pragma unroll
            for (int v = 0; v < V; ++v) tmp[v] = lds_out[(q + v) + ((q + v) >> PAD)];
            using Vec = __attribute__((ext_vector_type(4))) uint32_t;
            Vec w;
            __builtin_memcpy(&w, tmp, 16);
            *reinterpret_cast<Vec*>(out + q) = w;
        } else {
            for (int v = 0; v < V && q + v < total; ++v) out[q + v] = lds_out[(q + v) + ((q + v) >> PAD)];
        }
    }
}

template <int DT>
__global__ void __launch_bounds__(kMergeThreads)
k_merge_pass(const typename Elem<DT>::Storage* src, typename Elem<DT>::Storage* dst, int64_t n, int64_t run) {
    merge_tile_body<DT>(src, dst, n, run, int64_t(blockIdx.x) * merge_tile<DT>());
}

// ------------------------------------------------------------- host side --
struct LargeKernels {
    hipError_t (*chunk_sort)(const void* in, void* out, int64_t n, hipStream_t s) = nullptr;
    hipError_t (*merge_pass)(const void* src, void* dst, int64_t n, int64_t run, hipStream_t s) = nullptr;
    int chunk = 0;       // elements per chunk of the initial tile sort
    int tile = 0;        // elements per merge workgroup
};

template <int DT>
hipError_t large_chunk_sort(const void* in, void* out, int64_t n, hipStream_t s) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    constexpr int LW = chunk_logw<DT>();
    constexpr int LE = auto_loge(DT, LW);
    using C = TileCfg<E, LW, LE>;
    const int64_t chunk = int64_t(1) << LW;
    const int64_t full = n / chunk;
    const int64_t tail = n - full * chunk;
    if (full > 0) {
        hipLaunchKernelGGL((k_bitonic_tile<E, LW, LE>), dim3(unsigned((full + C::APB - 1) / C::APB)), dim3(C::BLOCK),
                           0, s, static_cast<const S*>(in), static_cast<S*>(out), full, int(chunk));
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (tail > 0) {
        hipLaunchKernelGGL((k_bitonic_tile<E, LW, LE>), dim3(1), dim3(C::BLOCK), 0, s,
                           static_cast<const S*>(in) + full * chunk, static_cast<S*>(out) + full * chunk, int64_t(1),
                           int(tail));
        return hipGetLastError();
    }
    return hipSuccess;
}

template <int DT>
hipError_t large_merge_pass(const void* src, void* dst, int64_t n, int64_t run, hipStream_t s) {
    using S = typename Elem<DT>::Storage;
    const int64_t tiles = (n + merge_tile<DT>() - 1) / merge_tile<DT>();
    if (tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_merge_pass<DT>, dim3(unsigned(tiles)), dim3(kMergeThreads), 0, s,
                       static_cast<const S*>(src), static_cast<S*>(dst), n, run);
    return hipGetLastError();
}

template <int DT>
LargeKernels large_kernels() {
    LargeKernels k;
    k.chunk_sort = &large_chunk_sort<DT>;
    k.merge_pass = &large_merge_pass<DT>;
    k.chunk = 1 << chunk_logw<DT>();
    k.tile = merge_tile<DT>();
    return k;
}

LargeKernels large_kernels_for(int dtype);

}  // namespace cs
