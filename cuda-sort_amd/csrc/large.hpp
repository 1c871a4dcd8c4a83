// Large-array sort: chunk sort + merge-path merge passes (the reference's
// SortMerge, cusort.hpp:2025-2300), and the sample-sort partition kernels
// (the reference's Sort / HybridSort partitioned path, cusort.hpp:1303-2023,
// 2488-2754), redesigned for gfx950.
//
// Merge pass.  Runs of length `run` are merged pairwise into runs of 2*run.
// Each workgroup (256 threads) produces one output tile of TILE = 256*IPT
// elements.  The co-rank of every tile start (how many elements of the left
// run precede its first output) is found beforehand by a separate kernel,
// one binary search per tile, so the dependent global loads of all tiles are
// in flight together instead of sitting at the start of every tile; a tile's
// end co-rank is its successor's start.  The merge workgroup then loads its
// two input slices into LDS (all loads of a thread in flight together), every
// thread finds its own co-rank inside LDS and merges IPT outputs into
// registers, and the tile leaves through the same LDS buffer as lane-striped
// 16-byte stores.
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
// Runs start at the chunk length, so every run pair holds whole merge tiles.
template <int DT> constexpr bool merge_tiles_align() { return (2 << chunk_logw<DT>()) % merge_tile<DT>() == 0; }

// Number of elements of A among the first d outputs of merge(A, B), ties
// taking A first: the smallest i with !(A[i] <= B[d-1-i]) in the feasible
// range [max(0, d-nb), min(d, na)].  Binary search by one thread; the upper
// levels of all the tiles' searches touch the same few lines, which stay in
// L2.
template <class E>
__device__ int64_t thread_corank(const typename E::Storage* A, int64_t na, const typename E::Storage* B, int64_t nb,
                                 int64_t d) {
    int64_t lo = d > nb ? d - nb : 0;
    int64_t hi = d < na ? d : na;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (!(E::key(B[d - 1 - mid]) < E::key(A[mid]))) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

constexpr int kCorankThreads = 256;

// Geometry of the merge tile that starts at output position g0 of an array
// of n elements with runs of length `run`.
struct MergeGeom {
    int64_t base;   // start of the run pair
    int64_t na, nb; // lengths of its left and right run
    int64_t d0, d1; // tile output range, relative to base
};

template <int TILE>
__device__ __forceinline__ MergeGeom merge_geom(int64_t n, int64_t run, int64_t g0) {
    MergeGeom g;
    g.base = (g0 / (2 * run)) * (2 * run);
    g.na = run < n - g.base ? run : n - g.base;
    const int64_t rem = n - g.base - g.na;
    g.nb = rem < 0 ? 0 : (rem < run ? rem : run);
    g.d0 = g0 - g.base;
    g.d1 = g.d0 + TILE < g.na + g.nb ? g.d0 + TILE : g.na + g.nb;
    return g;
}

// Same search inside the LDS staging buffer for one thread (binary); the
// left slice sits at staged positions [0, na), the right one at [na, na+nb),
// and position x is stored at word x + (x >> PIN).
template <class E, int PIN>
__device__ __forceinline__ int lds_corank(const typename E::Storage* lds, int na, int nb, int d) {
    int lo = d > nb ? d - nb : 0;
    int hi = d < na ? d : na;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int xb = na + d - 1 - mid;
        if (!(E::key(lds[xb + (xb >> PIN)]) < E::key(lds[mid + (mid >> PIN)]))) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One output tile [g0, g0 + TILE) of a merge pass over the array src[0, n)
// whose runs have length `run`, with the tile's co-ranks a0 / a1 (left-run
// elements before its first / one-past-last output); all kMergeThreads
// threads take part.
template <int DT>
__device__ __forceinline__ void merge_tile_body(const typename Elem<DT>::Storage* src,
                                                typename Elem<DT>::Storage* dst, int64_t n, int64_t run,
                                                int64_t g0, int64_t a0, int64_t a1) {
    using E = Elem<DT>;
    using S = typename E::Storage;
    constexpr int IPT = merge_ipt<DT>();
    constexpr int TILE = merge_tile<DT>();
    constexpr int V = 16 / int(sizeof(S)) > 0 ? 16 / int(sizeof(S)) : 1;
    // Staged inputs: one pad element per 2^PIN (threads start their serial
    // merges about IPT/2 elements apart in each slice, which would otherwise
    // put a half-wave on a few banks).  Outputs: one pad per 2^PAD.
    constexpr int PIN = sizeof(S) == 4 ? 3 : 2;
    constexpr int PAD = sizeof(S) == 4 ? 5 : (sizeof(S) == 8 ? 4 : 3);
    constexpr int LDS_IN = TILE + (TILE >> PIN);
    constexpr int LDS_OUT = TILE + (TILE >> PAD);
    __shared__ S lds[LDS_IN > LDS_OUT ? LDS_IN : LDS_OUT];

    if (g0 >= n) return;
    const MergeGeom g = merge_geom<TILE>(n, run, g0);
    const S* A = src + g.base;
    const S* B = A + g.na;
    const int tid = int(threadIdx.x);
    const int la = int(a1 - a0);
    const int lb = int((g.d1 - a1) - (g.d0 - a0));
    const int64_t b0 = g.d0 - a0;
    const int total = la + lb;
    // ---- both input slices into LDS (left slice first)
    #pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int idx = tid + k * kMergeThreads;
        if (idx < total) lds[idx + (idx >> PIN)] = idx < la ? A[a0 + idx] : B[b0 + (idx - la)];
    }
    __syncthreads();
    // ---- every thread merges IPT consecutive outputs into registers; one
    // unconditional (clamped) LDS read per output refills the side taken.
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
        int r = take_a ? i : la + j;   // staged position of the refill (unused once its side is exhausted)
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
    // ---- lane-striped stores
    S* out = dst + g.base + g.d0;
    const bool vec = (((uintptr_t)out) & 15) == 0 && V > 1;
    for (int q = tid * V; q < total; q += kMergeThreads * V) {
        S tmp[V];
        #pragma unroll
        for (int v = 0; v < V; ++v) {
            const int x = q + v < total ? q + v : q;
            tmp[v] = lds[x + (x >> PAD)];
        }
        if (q + V <= total) {
            if (vec) {   // 16-byte aligned tile start (plain merge passes)
                using Vec = __attribute__((ext_vector_type(4))) uint32_t;
                Vec w;
                __builtin_memcpy(&w, tmp, 16);
                *reinterpret_cast<Vec*>(out + q) = w;
            } else {     // tiles of buckets and requests start anywhere
                #pragma unroll
                for (int v = 0; v < V; ++v) out[q + v] = tmp[v];
            }
        } else {
            for (int v = 0; v < V && q + v < total; ++v) out[q + v] = tmp[v];
        }
    }
}

// Co-ranks of all tiles of a merge pass: cor[t] = co-rank of the first
// output of tile t.  The co-rank of a tile's one-past-last output is the next
// tile's cor (run pairs hold whole tiles: 2*run is a multiple of TILE), or
// the left run's length at the end of a pair.
template <int DT>
__global__ void __launch_bounds__(kCorankThreads)
k_merge_corank(const typename Elem<DT>::Storage* src, int64_t n, int64_t run, int64_t tiles, int64_t* cor) {
    using E = Elem<DT>;
    static_assert(merge_tiles_align<DT>(), "run pairs must hold whole merge tiles");
    constexpr int TILE = merge_tile<DT>();
    const int64_t t = int64_t(blockIdx.x) * kCorankThreads + threadIdx.x;
    if (t >= tiles) return;
    const MergeGeom g = merge_geom<TILE>(n, run, t * TILE);
    const typename E::Storage* A = src + g.base;
    cor[t] = thread_corank<E>(A, g.na, A + g.na, g.nb, g.d0);
}

// End co-rank of the tile with geometry g whose successor's start co-rank is
// next (read only when the tile does not end its run pair).
__device__ __forceinline__ int64_t merge_end_corank(const MergeGeom& g, const int64_t* next) {
    return g.d1 == g.na + g.nb ? g.na : *next;
}

template <int DT>
__global__ void __launch_bounds__(kMergeThreads)
k_merge_pass(const typename Elem<DT>::Storage* src, typename Elem<DT>::Storage* dst, int64_t n, int64_t run,
             const int64_t* cor) {
    const int64_t t = blockIdx.x;
    const int64_t g0 = t * merge_tile<DT>();
    const MergeGeom g = merge_geom<merge_tile<DT>()>(n, run, g0);
    merge_tile_body<DT>(src, dst, n, run, g0, cor[t], merge_end_corank(g, cor + t + 1));
}

// ------------------------------------------------------------- host side --
struct LargeKernels {
    hipError_t (*chunk_sort)(const void* in, void* out, int64_t n, hipStream_t s) = nullptr;
    // One merge pass; cor is device scratch of one int64 per merge tile.
    hipError_t (*merge_pass)(const void* src, void* dst, int64_t n, int64_t run, int64_t* cor,
                             hipStream_t s) = nullptr;
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
                           0, s, static_cast<const S*>(in), static_cast<S*>(out), full, int(chunk), 0u);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (tail > 0) {
        hipLaunchKernelGGL((k_bitonic_tile<E, LW, LE>), dim3(1), dim3(C::BLOCK), 0, s,
                           static_cast<const S*>(in) + full * chunk, static_cast<S*>(out) + full * chunk, int64_t(1),
                           int(tail), 0u);
        return hipGetLastError();
    }
    return hipSuccess;
}

template <int DT>
hipError_t large_merge_pass(const void* src, void* dst, int64_t n, int64_t run, int64_t* cor, hipStream_t s) {
    using S = typename Elem<DT>::Storage;
    const int64_t tiles = (n + merge_tile<DT>() - 1) / merge_tile<DT>();
    if (tiles <= 0) return hipSuccess;
    const int64_t cblocks = (tiles + kCorankThreads - 1) / kCorankThreads;
    hipLaunchKernelGGL(k_merge_corank<DT>, dim3(unsigned(cblocks)), dim3(kCorankThreads), 0, s,
                       static_cast<const S*>(src), n, run, tiles, cor);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_merge_pass<DT>, dim3(unsigned(tiles)), dim3(kMergeThreads), 0, s,
                       static_cast<const S*>(src), static_cast<S*>(dst), n, run, cor);
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
