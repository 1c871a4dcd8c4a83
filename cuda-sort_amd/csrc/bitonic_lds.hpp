// LDS-resident bitonic sort, one workgroup per array, any block size that is
// a multiple of 32 (partial last wave on wave64), run-time length.
//
// This serves the reference's block methods whose thread count is not a
// power-of-two register tiling: v1/v2 (one element per thread,
// cusort.hpp:608-1056), v3/v4 (any warp multiple, strided item loops,
// cusort.hpp:1061-1300, cusort_experiments.hpp:95-211), v5 (compile-time
// sizes, cusort_experiments.hpp:116-233) and the FULLSM switch for element
// types without a register tiling.  Every layer is one pass of
// compare-exchanges over W/2 disjoint pairs in LDS; the network, tie rule and
// padding are the reference's (reflection first layer i ^ (2^k - 1), then
// i ^ 2^j).
//
// One workgroup barrier follows every layer.  Running the layers whose
// pairs stay inside one wave's positions without it (the reference keeps
// sub-warp strides inside a warp, cusort.hpp:628-677, 712-737) was measured
// slower on gfx950: 4096 x 1024 int32 at 160 / 192 / 224 threads took 89 /
// 77 / 84 us against 82 / 69 / 75 us with a barrier after every layer.
//
// ASYNC (the reference's cooperative_groups::memcpy_async staging,
// cusort.hpp:808-837, 1021-1046): for key-only element types the array's
// bytes go from global memory straight into LDS by direct loads
// (global_load_lds_dword, one dword per lane, no VGPR round trip), and
// floating-point keys are then turned into their order-preserving integer
// keys in place.  Records keep the staged loads: their LDS layout splits the
// keys from the payloads, which a byte copy cannot do.  The store back goes
// through VGPRs either way (gfx950 has no LDS-to-global transfer).
//
// TESTSORTED: before the network and after every stage k >= 5 the block
// checks whether the array is already sorted and stops if so.  A sorted
// array is a fixed point of every remaining layer (no strict inversion ever
// exists), so the output is the same with or without the checks
// (cusort.hpp:394-458 checks at the same kind of checkpoints).
#pragma once

#include <type_traits>

#include "elem.hpp"

namespace cs {

template <class E>
__device__ __forceinline__ void lds_cx(typename E::Key* lk, int32_t* lp, int i, int j) {
    typename E::Key a = lk[i], b = lk[j];
    int32_t pa = 0, pb = 0;
    if constexpr (E::kPayload) { pa = lp[i]; pb = lp[j]; }
    if (b < a) {                        // strict inversion only: ties stay put
        lk[i] = b; lk[j] = a;
        if constexpr (E::kPayload) { lp[i] = pb; lp[j] = pa; }
    }
}

template <class E>
__device__ bool lds_sorted(const typename E::Key* lk, int n, int* flag) {
    if (threadIdx.x == 0) *flag = 1;
    __syncthreads();
    bool ok = true;
    for (int i = int(threadIdx.x) + 1; i < n; i += int(blockDim.x))
        if (lk[i] < lk[i - 1]) ok = false;
    if (!ok) *flag = 0;
    __syncthreads();
    const bool sorted = *flag != 0;
    __syncthreads();
    return sorted;
}

template <class E>
__global__ void __launch_bounds__(1024)
k_bitonic_lds(const typename E::Storage* in, typename E::Storage* out, int64_t arrays, int n, int logw,
              unsigned flags) {
    using Key = typename E::Key;
    extern __shared__ __align__(16) unsigned char smem[];
    const int W = 1 << logw;
    Key* lk = reinterpret_cast<Key*>(smem);
    int32_t* lp = reinterpret_cast<int32_t*>(smem + size_t(W) * sizeof(Key));
    __shared__ int sorted_flag;

    const int64_t a = blockIdx.x;
    if (a >= arrays) return;
    const typename E::Storage* src = in + a * int64_t(n);
    typename E::Storage* dst = out + a * int64_t(n);
    const int T = int(blockDim.x);
    const int tid = int(threadIdx.x);

    bool staged = true;
    if constexpr (!E::kPayload && sizeof(typename E::Storage) == sizeof(Key)) {
        if (flags & CS_FLAG_ASYNC) {   // workgroup-uniform
            staged = false;
            // dword d of the array to LDS dword d: each wave's lanes cover
            // consecutive dwords, so the wave's LDS base (its first dword)
            // is uniform, as the instruction requires
            const int dwords = int(size_t(n) * sizeof(Key) / 4);
            const uint32_t* gsrc = reinterpret_cast<const uint32_t*>(src);
            uint32_t* ldw = reinterpret_cast<uint32_t*>(lk);
            const int lane = tid & 63;
            for (int d = tid; d < dwords; d += T)
                __builtin_amdgcn_global_load_lds(gsrc + d, (__attribute__((address_space(3))) void*)(ldw + (d - lane)),
                                                 4, 0, 0);
            for (int i = n + tid; i < W; i += T) lk[i] = E::kPad;
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
            // keys that are not their own storage bits (fp totalOrder) are
            // converted in place, each thread its own positions
            if constexpr (std::is_same<E, Elem<CS_F32>>::value || std::is_same<E, Elem<CS_F64>>::value) {
                for (int i = tid; i < n; i += T) {
                    typename E::Storage raw;
                    __builtin_memcpy(&raw, &lk[i], sizeof(raw));
                    lk[i] = E::key(raw);
                }
            }
        }
    }
    if (staged) {
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
    }
    __syncthreads();

    const bool test = (flags & CS_FLAG_TESTSORTED) != 0;
    bool done = test && lds_sorted<E>(lk, n, &sorted_flag);
    for (int k = 1; k <= logw && !done; ++k) {
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
        if (test && k >= 5 && k < logw) done = lds_sorted<E>(lk, n, &sorted_flag);
    }

    for (int i = tid; i < n; i += T) dst[i] = E::make(lk[i], E::kPayload ? lp[i] : 0);
}

}  // namespace cs
