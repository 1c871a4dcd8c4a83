// Register-tile bitonic sort for gfx950 (wave64): one array per lane group
// of T = W / E lanes (several arrays per wave when T < 64, one workgroup per
// array when T > 64), E elements per lane held in VGPRs.
//
// Network: the all-ascending "reflection" bitonic network of the reference
// (cusort.hpp:545-547, 686-697; cusort_v6.hpp:38-128): stage k first
// compares i with i ^ (2^k - 1), then i with i ^ 2^j for j = k-2 .. 0, the
// lower index keeping the minimum and each side keeping its own element on
// equal keys.  What is new is how the network is mapped onto the hardware.
//
// Layouts.  The W = 2^w positions of an array are spread over register slot
// bits and lane bits.  Every layout used here is L(h): the e = log2(E) slot
// bits hold logical bits h-e+1 .. h, the lane bits hold the other logical
// bits in increasing order.  A compare-exchange on logical bit j is a pair
// of VGPRs (one v_min + one v_max per pair, no cross-lane traffic) whenever
// j is a slot bit of the current layout.  Stages k <= e run entirely in the
// identity layout L(e-1).  A stage k > e is cut into chunks of at most e bits
// from the top (the first chunk takes the remainder so the last chunk is
// always bits e-1..0), and the tile is re-laid between chunks by one LDS
// transpose (write every VGPR at its logical position, read back in the next
// layout).  So every one of the w(w+1)/2 layers costs one VALU op per element
// and the cross-lane work is moved to the LDS pipe: sum over k>e of
// ceil(k/e) transposes.
//
// The reflection of stage k > e is folded into the transpose that opens the
// stage: the upper half (bit k-1 = 1) is read back with logical bits
// 0..k-2 flipped, which turns the reflection into a plain compare on bit
// k-1; while the flip is in force the upper half merges descending (bit k-1
// is a slot bit there, so the direction is a compile-time property of each
// VGPR); the next transpose undoes the flip as it writes.  Position by
// position the result is the reference network's, so its padding and tie
// guarantees carry over: max-key padding placed at logical positions >= n
// stays there (it never moves to a lower logical index), records are never
// duplicated or lost.
#pragma once

#include <type_traits>

#include "elem.hpp"
#include "plan.hpp"

namespace cs {

template <int I, int N, class F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

// --------------------------------------------------------------- config --
template <class E, int LOGW, int LOGE> struct TileCfg {
    using Key = typename E::Key;
    using Storage = typename E::Storage;
    static constexpr int W = 1 << LOGW;
    static constexpr int NE = 1 << LOGE;             // elements per lane
    static constexpr int LOGT = LOGW - LOGE;
    static constexpr int T = 1 << LOGT;              // lanes per array
    static constexpr int BLOCK = T >= 256 ? T : 256; // threads per workgroup
    static constexpr int APB = BLOCK / T;            // arrays per workgroup
    static constexpr bool kMultiWave = T > 64;
    static constexpr int NTRANS = PlanOf<LOGW, LOGE>::value.ntrans;
    // LDS address of logical position x: x + (x >> PADSHIFT) keeps a
    // column of lanes 2^PADSHIFT elements apart on distinct banks.
    static constexpr int PADSHIFT = sizeof(Key) == 4 ? 5 : 4;
    static constexpr int REGION = W + (W >> PADSHIFT);
    static constexpr int NBUF = kMultiWave ? 2 : 1;
    static constexpr int LDS_KEYS = NTRANS ? NBUF * APB * REGION : 1;
    static constexpr int LDS_PAYLOAD = (NTRANS && E::kPayload) ? NBUF * APB * REGION : 1;
    static constexpr int LDS_BYTES = NTRANS ? LDS_KEYS * int(sizeof(Key)) + (E::kPayload ? LDS_PAYLOAD * 4 : 0) : 0;
    static_assert(LOGE <= LOGW, "more elements per lane than the array holds");
    static_assert(T <= 1024, "at most 1024 lanes per array");
    static_assert(LDS_BYTES <= 160 * 1024, "tile exceeds the 160 KB LDS of a gfx950 workgroup");
};

template <class E, int LOGW, int LOGE>
struct Tile {
    using C = TileCfg<E, LOGW, LOGE>;
    using Key = typename E::Key;
    static constexpr int NE = C::NE;
    static constexpr int e = LOGE;
    static constexpr int w = LOGW;

    // Elements whose compare-exchange is a compare and selects (records,
    // 8-byte keys), and every multi-wave tile, open with an odd-even merge
    // sort of each lane's slots: 191 compare-exchanges for 32 slots against
    // the bitonic stages' 240.  Single-wave tiles of 4-byte keys (one v_min +
    // one v_max per pair) keep the bitonic stages, whose layers carry more
    // independent pairs (measured faster there).
    static constexpr bool kLaneSort = E::kPayload || sizeof(Key) > 4 || C::T > 64;
    using P = PlanOf<LOGW, LOGE, kLaneSort>;

    Key k[NE];
    int32_t p[NE];

    template <bool kDesc>
    __device__ __forceinline__ void cx(int a, int b) { cmpx<E, kDesc>(k[a], p[a], k[b], p[b]); }

    template <int I>
    __device__ __forceinline__ void run(int t, Key* lk, int32_t* lp) {
        if constexpr (I < P::value.nops) {
            constexpr Op op = P::value.ops[I];
            if constexpr (op.kind == OP_LANESORT) {
                sfor<0, OemOf<NE>::value.n>([&](auto i_) {
                    constexpr int i = decltype(i_)::value;
                    cx<false>(OemOf<NE>::value.c[i].a, OemOf<NE>::value.c[i].b);
                });
            } else if constexpr (op.kind == OP_REFLECT) {
                constexpr int mask = (2 << op.bit) - 1;
                sfor<0, NE>([&](auto r_) {
                    constexpr int r = decltype(r_)::value;
                    if constexpr (((r >> op.bit) & 1) == 0) cx<false>(r, r ^ mask);
                });
            } else if constexpr (op.kind == OP_CE) {
                constexpr int s = op.bit - (op.cur_h - e + 1);
                static_assert(s >= 0 && s < e, "compare on a lane bit");
                constexpr int sd = op.dirbit >= 0 ? op.dirbit - (op.cur_h - e + 1) : -1;
                sfor<0, NE>([&](auto r_) {
                    constexpr int r = decltype(r_)::value;
                    if constexpr (((r >> s) & 1) == 0) {
                        constexpr bool desc = sd >= 0 && ((r >> (sd >= 0 ? sd : 0)) & 1);
                        cx<desc>(r, r | (1 << s));
                    }
                });
            } else {
                transpose<op.from_h, op.to_h, op.flip_read, op.flip_write, op.index>(t, lk, lp);
            }
            run<I + 1>(t, lk, lp);
        }
    }

    template <int FH, int TH, int FR, int FW, int IDX>
    __device__ __forceinline__ void transpose(int t, Key* lk, int32_t* lp) {
        static_assert(FW < 0 || FW == FH, "write flip must be the source layout's top bit");
        static_assert(FR < 0 || FR == TH, "read flip must be the target layout's top bit");
        constexpr int PS = C::PADSHIFT;
        constexpr int buf = C::kMultiWave ? (IDX & 1) * C::APB * C::REGION : 0;
        // Lane parts of the addresses (runtime, shared by all slots); the
        // slot parts are compile-time offsets (lds_addr is additive over
        // disjoint bits).
        const int wa = lds_addr(transpose_base(t, FH, e, false), PS) + buf;
        const int ra = lds_addr(transpose_base(t, TH, e, false), PS) + buf;
        const int waf = FW >= 0 ? lds_addr(transpose_base(t, FH, e, true), PS) + buf : 0;
        const int raf = FR >= 0 ? lds_addr(transpose_base(t, TH, e, true), PS) + buf : 0;
        sfor<0, NE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            constexpr bool flip = slot_flipped(r, e, FW >= 0);
            constexpr int off = lds_addr(transpose_slot(r, FH, e, FW >= 0), PS);
            const int a = (flip ? waf : wa) + off;
            lk[a] = k[r];
            if constexpr (E::kPayload) lp[a] = p[r];
        });
        if constexpr (C::kMultiWave) {
            __syncthreads();
        } else {
            // One wave owns the region: DS instructions of a wave execute in
            // order, so only the compiler has to be kept from reordering.
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        sfor<0, NE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            constexpr bool flip = slot_flipped(r, e, FR >= 0);
            constexpr int off = lds_addr(transpose_slot(r, TH, e, FR >= 0), PS);
            const int a = (flip ? raf : ra) + off;
            k[r] = lk[a];
            if constexpr (E::kPayload) p[r] = lp[a];
        });
        if constexpr (!C::kMultiWave) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }

    // Lane-striped I/O order: slot r of lane t is array position
    // ((r / VS) * T + t) * VS + r % VS, so for every group of VS slots the
    // lanes of an array cover consecutive 16-byte chunks (coalesced global
    // accesses).  io_transpose moves the tile between that order and the
    // identity layout L(e-1) (slot r of lane t = position t*NE + r) through
    // LDS; both sides are additive lane/slot address splits.
    template <int VS, bool kToIdentity>
    __device__ __forceinline__ void io_transpose(int t, Key* lk, int32_t* lp, int buf) {
        constexpr int PS = C::PADSHIFT;
        const int sa = lds_addr(t * VS, PS) + buf;
        const int ia = lds_addr(t * NE, PS) + buf;
        sfor<0, NE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            constexpr int soff = lds_addr((r / VS) * C::T * VS + (r % VS), PS);
            constexpr int ioff = lds_addr(r, PS);
            const int a = kToIdentity ? sa + soff : ia + ioff;
            lk[a] = k[r];
            if constexpr (E::kPayload) lp[a] = p[r];
        });
        if constexpr (C::kMultiWave) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        sfor<0, NE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            constexpr int soff = lds_addr((r / VS) * C::T * VS + (r % VS), PS);
            constexpr int ioff = lds_addr(r, PS);
            const int a = kToIdentity ? ia + ioff : sa + soff;
            k[r] = lk[a];
            if constexpr (E::kPayload) p[r] = lp[a];
        });
        if constexpr (!C::kMultiWave) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
};

// TESTSORTED (the reference's TestIfSorted, cusort_v6.hpp:103-142): is the
// freshly loaded array already non-decreasing?  The tile is still in the
// lane-striped load order (slot q*VS + j of lane t = position (q*T + t)*VS
// + j; positions >= valid hold the maximum key, so a sorted valid prefix
// stays sorted).  Each lane checks its groups in registers; the first key
// of every group goes through the group's LDS region so that lane t can
// compare its group ends with the next positions (for t = T-1 that is
// lane 0's next group).  Group-uniform result; a tile spanning several
// waves makes it uniform over the workgroup, whose network barriers need
// every lane.  A sorted array is a fixed point of the network (no strict
// inversion), so skipping the network leaves the output unchanged.
template <class E, int LOGW, int LOGE, int VS>
__device__ __forceinline__ bool tile_striped_sorted(const typename E::Key* k, int t, typename E::Key* lk) {
    using C = TileCfg<E, LOGW, LOGE>;
    using Key = typename E::Key;
    constexpr int NE = C::NE, T = C::T, NG = NE / VS;
    bool ok = true;
    sfor<0, NG>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        sfor<0, VS - 1>([&](auto j_) {
            constexpr int j = decltype(j_)::value;
            ok = ok && !(k[q * VS + j + 1] < k[q * VS + j]);
        });
    });
    if constexpr (T == 1) {
        sfor<0, NG - 1>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            ok = ok && !(k[(q + 1) * VS] < k[q * VS + VS - 1]);
        });
        return ok;
    } else {
        auto sync = [] {
            if constexpr (C::kMultiWave) {
                __syncthreads();
            } else {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        };
        constexpr int FLAG = NG * T;   // one word past the group-first keys
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            lk[q * T + t] = k[q * VS];
        });
        if (t == 0) lk[FLAG] = Key(1);
        sync();
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            const int x = q * T + t + 1;
            const Key nx = lk[x < FLAG ? x : 0];
            ok = ok && (x >= FLAG || !(nx < k[q * VS + VS - 1]));
        });
        if (!ok) lk[FLAG] = Key(0);
        sync();
        const bool sorted = lk[FLAG] != Key(0);
        sync();   // the region is reused by the network
        if constexpr (C::kMultiWave && C::APB > 1) return __syncthreads_and(sorted ? 1 : 0) != 0;
        return sorted;
    }
}

// ------------------------------------------------------------ the kernel --
// Global I/O is lane-striped (coalesced 16-byte accesses).  Keys without
// payload are loaded straight into the registers in that order and the
// order is simply declared to be the identity layout: the sorted output
// does not depend on the input order, and max-key padding cannot enter the
// valid prefix of a key-only sort.  Records must start with logical
// position = input position (so the padding guarantee applies), so their
// striped load is re-laid to the identity layout by one LDS transpose.  The
// sorted tile leaves the identity layout through one more LDS transpose so
// that the stores are lane-striped too.
// Sorts one array of `valid` elements (valid <= W; 0 for an idle group)
// from src into dst with the T lanes t = 0..T-1 of this group; lk / lp are
// the group's LDS region (C::REGION words per buffer).  Every lane of the
// workgroup must call it when C::kMultiWave (it contains barriers).
// With test (TESTSORTED) an array found sorted right after the load is
// stored unchanged without running the network.
template <class E, int LOGW, int LOGE>
__device__ __forceinline__ void tile_sort(const typename E::Storage* src, typename E::Storage* dst, int valid,
                                          int t, typename E::Key* lk, int32_t* lp, bool test = false) {
    using C = TileCfg<E, LOGW, LOGE>;
    using Storage = typename E::Storage;
    constexpr int NE = C::NE;
    constexpr int T = C::T;
    constexpr int V = 16 / int(sizeof(Storage)) > 0 ? 16 / int(sizeof(Storage)) : 1;   // elements per 16 B
    constexpr int VS = V < NE ? V : NE;                                               // per striped group
    using Vec = __attribute__((ext_vector_type(4))) uint32_t;
    constexpr int kBufLast = C::kMultiWave ? (C::NTRANS & 1) * C::APB * C::REGION : 0;
    constexpr int kBufFirst = C::kMultiWave ? C::APB * C::REGION : 0;

    Tile<E, LOGW, LOGE> tile;
    // 16-byte accesses of V elements.  Global memory needs only dword
    // alignment for them, which every element has, so arrays that start
    // anywhere (ragged segments, partition buckets) use them too; they are
    // written as memcpy of an element pointer, so the compiler sees the
    // element's alignment, not the vector's.
    constexpr bool vec_in = VS == V;
    constexpr bool vec_out = VS == V;

    // ---- load, lane-striped.  When every group of this lane is in range
    // (all lanes of a full array), the loads form one branch-free block, all
    // in flight together: 16-byte vectors (element loads for the 16-byte
    // record tilings whose groups are narrower than a vector).  A lane
    // with a group past the end bounds-checks each group.  Stores are split
    // the same way.
    constexpr int NG = NE / VS;
    const bool full = ((NG - 1) * T + t) * VS + VS <= valid;   // every group of this lane in range
    const bool full_in = vec_in && full;
    const bool full_out = vec_out && full;
    if (full_in) {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            Vec v;
            __builtin_memcpy(&v, src + (q * T + t) * VS, 16);
            Storage tmp[V];
            __builtin_memcpy(tmp, &v, 16);
            sfor<0, VS>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                tile.k[q * VS + j] = E::key(tmp[j]);
                tile.p[q * VS + j] = E::payload(tmp[j]);
            });
        });
    } else if (full) {   // in range but not 16-byte aligned: element loads, no branches
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            sfor<0, VS>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                const Storage s = src[(q * T + t) * VS + j];
                tile.k[q * VS + j] = E::key(s);
                tile.p[q * VS + j] = E::payload(s);
            });
        });
    } else {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            const int g = (q * T + t) * VS;
            if (vec_in && g + VS <= valid) {
                Vec v;
                __builtin_memcpy(&v, src + g, 16);
                Storage tmp[V];
                __builtin_memcpy(tmp, &v, 16);
                sfor<0, VS>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    tile.k[q * VS + j] = E::key(tmp[j]);
                    tile.p[q * VS + j] = E::payload(tmp[j]);
                });
            } else {
                sfor<0, VS>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    if (g + j < valid) {
                        const Storage s = src[g + j];
                        tile.k[q * VS + j] = E::key(s);
                        tile.p[q * VS + j] = E::payload(s);
                    } else {
                        tile.k[q * VS + j] = E::kPad;
                        tile.p[q * VS + j] = INT32_MAX;
                    }
                });
            }
        });
    }
    // ---- network (skipped for an array found sorted under TESTSORTED: the
    // tile is then still in the striped order the stores use)
    if (!(test && tile_striped_sorted<E, LOGW, LOGE, VS>(tile.k, t, lk))) {
        if constexpr (E::kPayload && C::NTRANS > 0) tile.template io_transpose<VS, true>(t, lk, lp, kBufFirst);
        tile.template run<0>(t, lk, lp);
        if constexpr (C::NTRANS > 0) tile.template io_transpose<VS, false>(t, lk, lp, kBufLast);
    }

    // ---- store, lane-striped
    auto store_group = [&](auto q_) {
        constexpr int q = decltype(q_)::value;
        Storage tmp[V];
        sfor<0, VS>([&](auto j_) {
            constexpr int j = decltype(j_)::value;
            tmp[j] = E::make(tile.k[q * VS + j], tile.p[q * VS + j]);
        });
        Vec v;
        __builtin_memcpy(&v, tmp, 16);
        __builtin_memcpy(dst + (q * T + t) * VS, &v, 16);
    };
    if (full_out) {
        sfor<0, NG>(store_group);
    } else if (full) {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            sfor<0, VS>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                dst[(q * T + t) * VS + j] = E::make(tile.k[q * VS + j], tile.p[q * VS + j]);
            });
        });
    } else {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            const int g = (q * T + t) * VS;
            if (vec_out && g + VS <= valid) {
                store_group(q_);
            } else {
                sfor<0, VS>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    if (g + j < valid) dst[g + j] = E::make(tile.k[q * VS + j], tile.p[q * VS + j]);
                });
            }
        });
    }
}

// Batched launch: array a = in + a*n, APB arrays per workgroup.
template <class E, int LOGW, int LOGE>
__global__ void __launch_bounds__((TileCfg<E, LOGW, LOGE>::BLOCK))
k_bitonic_tile(const typename E::Storage* in, typename E::Storage* out, int64_t arrays, int n, unsigned flags) {
    using C = TileCfg<E, LOGW, LOGE>;
    __shared__ typename E::Key lds_k[C::LDS_KEYS];
    __shared__ int32_t lds_p[C::LDS_PAYLOAD];
    const int tid = int(threadIdx.x);
    const int group = tid >> C::LOGT;
    const int t = tid & (C::T - 1);
    const int64_t a = int64_t(blockIdx.x) * C::APB + group;
    const bool active = a < arrays;
    tile_sort<E, LOGW, LOGE>(in + (active ? a : 0) * int64_t(n), out + (active ? a : 0) * int64_t(n),
                             active ? n : 0, t, lds_k + (C::NTRANS ? group * C::REGION : 0),
                             lds_p + ((C::NTRANS && E::kPayload) ? group * C::REGION : 0),
                             (flags & CS_FLAG_TESTSORTED) != 0);
}

}  // namespace cs
