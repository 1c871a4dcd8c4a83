// Cross-lane register bitonic sort for 4-byte keys without payload (int32,
// uint32, fp32 through its totalOrder key), W = 2^LOGW positions per array,
// 32 <= W <= 2048: every array is held by T = W / 32 lanes of one wave with
// 32 keys per lane in VGPRs (64 / T arrays per wave).
//
// Network: an odd-even merge sorter on each lane's 32 registers (191
// comparators instead of the bitonic sorter's 240), then the bitonic merge
// stages k = 6 .. LOGW of the standard (direction) formulation: stage k
// compares i with i ^ 2^j for j = k-1 .. 0, ascending where logical bit k of
// i is 0.  Every compare-exchange goes through a median of three: with c the
// smallest key, med3(a, b, c) = min(a, b), with c the largest it is max(a, b),
// so one v_med3 per element applies a direction held per lane in a VGPR, and
// the direction costs nothing.
//
// Placement of the logical bits: bits 0..4 are the register index, bits 5..
// are lane bits at the physical positions kPos (xlane_plan.hpp).
//  * A layer on a register bit is one v_med3 (or v_min / v_max) per element.
//  * A layer on a lane bit at position 0, 1 or 3 is one DPP move of the
//    partner's value (quad_perm / row_ror:8) plus one v_med3 whose constant
//    also encodes which side of the pair the lane is on: two VALU operations
//    per element and no LDS traffic.
//  * A layer on position 4 or 5 swaps register pairs across the two rows /
//    halves with v_permlane16_swap / v_permlane32_swap (two VALU issue slots
//    each), computes both results of the pair with two v_med3 and swaps back:
//    three slots per element.
//  * Stages k >= LDSK instead move through an LDS "window": the tile is
//    written in the standard layout (16-byte stores, lane t holds positions
//    32 t .. 32 t + 31) and read back with logical bits k-5 .. k-1 in the
//    registers, so the stage's top five layers are register layers; the tile
//    then returns to the standard layout for the remaining layers.
// Keys carry no payload, so ties need no rule: any sorting network yields the
// same key sequence, and max-key padding at positions >= n (equal to or
// larger than every valid key) leaves the valid prefix intact.  Key-only
// input order is irrelevant, so the lane-striped 16-byte loads are declared
// to be the standard layout; the sorted tile leaves through one LDS pass so
// that the stores are lane-striped 16-byte stores too.
#pragma once

#include <type_traits>

#include "bitonic_tile.hpp"
#include "xlane_plan.hpp"

namespace cs {

template <class K> struct KeyLimits;
template <> struct KeyLimits<int32_t> {
    static constexpr int32_t lo = INT32_MIN;
    static constexpr int32_t hi = INT32_MAX;
};
template <> struct KeyLimits<uint32_t> {
    static constexpr uint32_t lo = 0u;
    static constexpr uint32_t hi = 0xFFFFFFFFu;
};

template <class K> __device__ __forceinline__ K kmin(K a, K b) { return __builtin_elementwise_min(a, b); }
template <class K> __device__ __forceinline__ K kmax(K a, K b) { return __builtin_elementwise_max(a, b); }
// Median of three (one v_med3_i32 / v_med3_u32).
template <class K> __device__ __forceinline__ K med3(K a, K b, K c) { return kmax(kmin(a, b), kmin(kmax(a, b), c)); }

template <int CTRL, class K>
__device__ __forceinline__ K dpp_mov(K v) {
    return __builtin_bit_cast(K, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

// Rows-swap of two registers: position 4 pairs lane l with l ^ 16, position 5
// with l ^ 32.  Afterwards, in every lane, .first holds the value of the pair
// member whose bit is 0 and .second that of the member whose bit is 1 (of
// register a in the lanes whose bit is 0, of register b in the others).
template <int POS, class K>
__device__ __forceinline__ void rows_swap(K& a, K& b) {
    static_assert(POS == 4 || POS == 5, "rows_swap crosses lane bit 4 or 5");
    const uint32_t x = __builtin_bit_cast(uint32_t, a), y = __builtin_bit_cast(uint32_t, b);
    if constexpr (POS == 4) {
        const auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
        a = __builtin_bit_cast(K, uint32_t(r[0]));
        b = __builtin_bit_cast(K, uint32_t(r[1]));
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
        a = __builtin_bit_cast(K, uint32_t(r[0]));
        b = __builtin_bit_cast(K, uint32_t(r[1]));
    }
}

__device__ __forceinline__ void wave_lds_fence() {
    // One wave owns every region it touches: DS instructions of a wave are
    // executed in order, so only compiler reordering has to be prevented.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class E, int LOGW, int LDSK>
struct XlTile {
    using K = typename E::Key;
    static_assert(!E::kPayload && sizeof(K) == 4, "cross-lane engine: 4-byte keys without payload");
    static_assert(LOGW >= 5 && LOGW <= 11, "cross-lane engine: 32..2048 positions per array");
    // With 64 lanes per array the sixth lane bit sits at position 2 (lane ^ 4,
    // no single DPP move): only a window stage may touch it.
    static_assert(LOGW <= 10 || LDSK <= 11, "2048 positions need the window for stage 11");
    static constexpr int LOGT = LOGW - xl::kLogE;
    static constexpr int T = 1 << LOGT;
    static constexpr K kLo = KeyLimits<K>::lo;
    static constexpr K kHi = KeyLimits<K>::hi;
    using V4 = __attribute__((ext_vector_type(4))) K;

    K k[xl::kE];

    // ---------------------------------------------------- compare-exchange
    template <int A, int B, bool kAsc>
    __device__ __forceinline__ void cx(K cl) {
        const K x = k[A], y = k[B];
        if constexpr (kAsc) {
            k[A] = kmin(x, y);
            k[B] = kmax(x, y);
        } else {
            k[A] = med3(x, y, cl);
            k[B] = med3(x, y, K(~cl));
        }
    }

    // In-lane odd-even merge sort, ascending (kAsc) or in the direction cl.
    template <bool kAsc>
    __device__ __forceinline__ void lane_sort(K cl) {
        sfor<0, xl::kOem32.n>([&](auto i_) {
            constexpr int i = decltype(i_)::value;
            cx<xl::kOem32.c[i].a, xl::kOem32.c[i].b, kAsc>(cl);
        });
    }

    // Layer on register bit J.
    template <int J, bool kAsc>
    __device__ __forceinline__ void reg_layer(K cl) {
        sfor<0, xl::kE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            if constexpr (((r >> J) & 1) == 0) cx<r, r | (1 << J), kAsc>(cl);
        });
    }
    // Register layers on bits HI, HI-1, ..., 0.
    template <int HI, bool kAsc>
    __device__ __forceinline__ void reg_layers_down(K cl) {
        if constexpr (HI >= 0) {
            reg_layer<HI, kAsc>(cl);
            reg_layers_down<HI - 1, kAsc>(cl);
        }
    }

    // Layer on logical lane bit L (logical bit 5 + L), lane t.
    template <int L>
    __device__ __forceinline__ void lane_layer(K cl, int t) {
        constexpr int pos = xl::kPos[L];
        if constexpr (xl::dpp_pos(pos)) {
            constexpr int ctrl = pos == 0 ? 0xB1 : (pos == 1 ? 0x4E : 0x128);   // ^1, ^2 (quad_perm), ^8 (row_ror:8)
            const K c = ((t >> L) & 1) ? K(~cl) : cl;   // the upper member takes the other choice
            // All partner values first, then the medians: distinct temporaries
            // keep every DPP read clear of the previous VALU write.
            K p[xl::kE];
            sfor<0, xl::kE>([&](auto r_) {
                constexpr int r = decltype(r_)::value;
                p[r] = dpp_mov<ctrl>(k[r]);
            });
            sfor<0, xl::kE>([&](auto r_) {
                constexpr int r = decltype(r_)::value;
                k[r] = med3(k[r], p[r], c);
            });
        } else {
            sfor<0, xl::kE / 2>([&](auto h_) {
                constexpr int r = 2 * decltype(h_)::value;
                K a = k[r], b = k[r + 1];
                rows_swap<pos>(a, b);                    // (lower member, upper member) of one register's pair
                K lo = med3(a, b, cl), hi = med3(a, b, K(~cl));
                rows_swap<pos>(lo, hi);                  // back: new value of register r, r + 1 in every lane
                k[r] = lo;
                k[r + 1] = hi;
            });
        }
    }

    // ------------------------------------------------------------- LDS
    __device__ __forceinline__ void put_std(K* lk, int t) {
        K* base = lk + xl::lds_addr(t << xl::kLogE);
        sfor<0, xl::kE / 4>([&](auto j_) {
            constexpr int j = decltype(j_)::value;
            *reinterpret_cast<V4*>(base + 4 * j) = V4{k[4 * j], k[4 * j + 1], k[4 * j + 2], k[4 * j + 3]};
        });
    }
    __device__ __forceinline__ void get_std(const K* lk, int t) {
        const K* base = lk + xl::lds_addr(t << xl::kLogE);
        sfor<0, xl::kE / 4>([&](auto j_) {
            constexpr int j = decltype(j_)::value;
            const V4 v = *reinterpret_cast<const V4*>(base + 4 * j);
            __builtin_memcpy(&k[4 * j], &v, 16);
        });
    }
    template <int KS>
    __device__ __forceinline__ void get_win(const K* lk, int u) {
        const K* base = lk + xl::lds_addr(xl::win_lane_pos(u, KS));
        sfor<0, xl::kE>([&](auto s_) {
            constexpr int s = decltype(s_)::value;
            k[s] = base[xl::lds_addr(xl::win_slot_pos(s, KS))];
        });
    }
    template <int KS>
    __device__ __forceinline__ void put_win(K* lk, int u) {
        K* base = lk + xl::lds_addr(xl::win_lane_pos(u, KS));
        sfor<0, xl::kE>([&](auto s_) {
            constexpr int s = decltype(s_)::value;
            base[xl::lds_addr(xl::win_slot_pos(s, KS))] = k[s];
        });
    }

    // ---------------------------------------------------- merge stages
    template <int KS>
    __device__ __forceinline__ void stages(int t, K* lk) {
        if constexpr (KS <= LOGW) {
            constexpr bool last = KS == LOGW;
            // Direction of stage KS: logical bit KS, which is lane bit KS - 5
            // in the standard layout and in the window layout alike.
            const K cl = (!last && ((t >> (KS - 5)) & 1)) ? kHi : kLo;
            if constexpr (KS < LDSK) {
                sfor<0, KS - 5>([&](auto i_) { lane_layer<KS - 6 - decltype(i_)::value>(cl, t); });
                reg_layers_down<4, last>(cl);
            } else {
                put_std(lk, t);
                wave_lds_fence();
                get_win<KS>(lk, t);
                wave_lds_fence();
                reg_layers_down<4, last>(cl);        // logical bits KS-1 .. KS-5
                put_win<KS>(lk, t);
                wave_lds_fence();
                get_std(lk, t);
                wave_lds_fence();
                // logical bits KS-6 .. 5 (lane bits), then 4 .. 0
                sfor<0, (KS > 10 ? KS - 10 : 0)>([&](auto i_) { lane_layer<KS - 11 - decltype(i_)::value>(cl, t); });
                reg_layers_down<(KS - 6 < 4 ? KS - 6 : 4), last>(cl);
            }
            stages<KS + 1>(t, lk);
        }
    }
};

// Load of one array slot: group q of lane t is array positions
// 4 (q T + t) .. + 3, kept as raw storage bits (no use of the data, so the
// loads stay in flight); positions >= valid hold the padding element.
template <class E, int LOGW>
__device__ __forceinline__ void xl_load(typename E::Storage (&raw)[xl::kE], const typename E::Storage* src,
                                        int valid, int t) {
    using S = typename E::Storage;
    constexpr int T = 1 << (LOGW - xl::kLogE);
    constexpr int NG = xl::kE / 4;   // 16-byte groups per lane
    using Vec = __attribute__((ext_vector_type(4))) uint32_t;
    static_assert(sizeof(S) == 4, "4-byte storage");
    const bool full = ((NG - 1) * T + t) * 4 + 4 <= valid;
    if (full && ((uintptr_t)src & 15) == 0) {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            // Copied out whole: a bit cast of a vector component (v.y) is
            // lowered as a load of the first component.
            const Vec v = *reinterpret_cast<const Vec*>(src + (q * T + t) * 4);
            __builtin_memcpy(&raw[4 * q], &v, 16);
        });
    } else if (full) {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            sfor<0, 4>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                raw[4 * q + j] = src[(q * T + t) * 4 + j];
            });
        });
    } else {
        const S pad = E::make(E::kPad, 0);
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            const int g = (q * T + t) * 4;
            sfor<0, 4>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                raw[4 * q + j] = g + j < valid ? src[g + j] : pad;
            });
        });
    }
}

// Segment form of xl_load (ragged segments and partition buckets start
// anywhere and mostly fill part of the tile).  The batched kernel keeps
// xl_load: with these paths in it the 4096 x 1024 int32 batch measured
// 0.8 % slower.
template <class E, int LOGW>
__device__ __forceinline__ void xl_load_seg(typename E::Storage (&raw)[xl::kE], const typename E::Storage* src,
                                        int valid, int t) {
    using S = typename E::Storage;
    constexpr int T = 1 << (LOGW - xl::kLogE);
    constexpr int NG = xl::kE / 4;   // 16-byte groups per lane
    using Vec = __attribute__((ext_vector_type(4))) uint32_t;
    static_assert(sizeof(S) == 4, "4-byte storage");
    const bool full = ((NG - 1) * T + t) * 4 + 4 <= valid;
    // 16-byte loads need only dword alignment in global memory: memcpy from
    // the element pointer tells the compiler exactly that, so segments that
    // start anywhere load by vectors too, and a partial lane loads its
    // in-range groups by vectors.
    if (full && ((uintptr_t)src & 15) == 0) {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            // Copied out whole: a bit cast of a vector component (v.y) is
            // lowered as a load of the first component.
            const Vec v = *reinterpret_cast<const Vec*>(src + (q * T + t) * 4);
            __builtin_memcpy(&raw[4 * q], &v, 16);
        });
    } else if (full) {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            Vec v;
            __builtin_memcpy(&v, src + (q * T + t) * 4, 16);
            __builtin_memcpy(&raw[4 * q], &v, 16);
        });
    } else {
        const S pad = E::make(E::kPad, 0);
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            const int g = (q * T + t) * 4;
            if (g + 4 <= valid) {
                Vec v;
                __builtin_memcpy(&v, src + g, 16);
                __builtin_memcpy(&raw[4 * q], &v, 16);
            } else {
                sfor<0, 4>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    raw[4 * q + j] = g + j < valid ? src[g + j] : pad;
                });
            }
        });
    }
}

// Sorts the loaded slot: key transform (positions >= valid become padding),
// network, and one LDS pass that leaves the sorted keys in striped order:
// out[4 q + j] is array position 4 (q T + t) + j.
template <class E, int LOGW, int LDSK>
__device__ __forceinline__ void xl_network(const typename E::Storage (&raw)[xl::kE],
                                           typename E::Key (&out)[xl::kE], int valid, int t, typename E::Key* lk) {
    using Tile = XlTile<E, LOGW, LDSK>;
    constexpr int T = Tile::T;
    constexpr int NG = xl::kE / 4;

    Tile tile;
    if (((NG - 1) * T + t) * 4 + 4 <= valid) {
        sfor<0, xl::kE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            tile.k[r] = E::key(raw[r]);
        });
    } else {
        sfor<0, xl::kE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            tile.k[r] = ((r / 4) * T + t) * 4 + r % 4 < valid ? E::key(raw[r]) : E::kPad;
        });
    }
    // ---- in-lane sort: ascending where logical bit 5 (lane bit 0) is 0
    if constexpr (LOGW == xl::kLogE) {
        tile.template lane_sort<true>(Tile::kLo);
    } else {
        tile.template lane_sort<false>((t & 1) ? Tile::kHi : Tile::kLo);
    }
    // ---- merge stages
    tile.template stages<xl::kLogE + 1>(t, lk);
    // ---- to striped order through LDS
    tile.put_std(lk, t);
    wave_lds_fence();
    sfor<0, NG>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        const typename Tile::V4 v = *reinterpret_cast<const typename Tile::V4*>(lk + xl::lds_addr((q * T + t) * 4));
        __builtin_memcpy(&out[4 * q], &v, 16);
    });
    wave_lds_fence();   // the region is reused by the next tile
}

// Stores of a sorted slot through a plain pointer (any alignment, any valid).
template <class E, int LOGW>
__device__ __forceinline__ void xl_store(const typename E::Key (&out)[xl::kE], typename E::Storage* dst, int valid,
                                         int t) {
    using S = typename E::Storage;
    constexpr int T = 1 << (LOGW - xl::kLogE);
    constexpr int NG = xl::kE / 4;
    using Vec = __attribute__((ext_vector_type(4))) uint32_t;
    const bool full = ((NG - 1) * T + t) * 4 + 4 <= valid;
    if (full && ((uintptr_t)dst & 15) == 0) {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            S e[4];
            sfor<0, 4>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                e[j] = E::make(out[4 * q + j], 0);
            });
            Vec v;
            __builtin_memcpy(&v, e, 16);
            *reinterpret_cast<Vec*>(dst + (q * T + t) * 4) = v;
        });
    } else {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            const int g = (q * T + t) * 4;
            sfor<0, 4>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                if (full || g + j < valid) dst[g + j] = E::make(out[4 * q + j], 0);
            });
        });
    }
}

// Segment form of xl_store: 16-byte stores at any dword alignment, and by
// groups in a partial lane.
template <class E, int LOGW>
__device__ __forceinline__ void xl_store_seg(const typename E::Key (&out)[xl::kE], typename E::Storage* dst, int valid,
                                         int t) {
    using S = typename E::Storage;
    constexpr int T = 1 << (LOGW - xl::kLogE);
    constexpr int NG = xl::kE / 4;
    using Vec = __attribute__((ext_vector_type(4))) uint32_t;
    const bool full = ((NG - 1) * T + t) * 4 + 4 <= valid;
    if (full && ((uintptr_t)dst & 15) == 0) {   // 16-byte aligned arrays: their own branch, as for the loads
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            S e[4];
            sfor<0, 4>([&](auto j_) {
                constexpr int j = decltype(j_)::value;
                e[j] = E::make(out[4 * q + j], 0);
            });
            Vec v;
            __builtin_memcpy(&v, e, 16);
            *reinterpret_cast<Vec*>(dst + (q * T + t) * 4) = v;
        });
        return;
    }
    auto store_group = [&](auto q_) {
        constexpr int q = decltype(q_)::value;
        S e[4];
        sfor<0, 4>([&](auto j_) {
            constexpr int j = decltype(j_)::value;
            e[j] = E::make(out[4 * q + j], 0);
        });
        __builtin_memcpy(dst + (q * T + t) * 4, e, 16);   // dword alignment suffices
    };
    if (full) {
        sfor<0, NG>(store_group);
    } else {
        sfor<0, NG>([&](auto q_) {
            constexpr int q = decltype(q_)::value;
            const int g = (q * T + t) * 4;
            if (g + 4 <= valid) {
                store_group(q_);
            } else {
                sfor<0, 4>([&](auto j_) {
                    constexpr int j = decltype(j_)::value;
                    if (g + j < valid) dst[g + j] = E::make(out[4 * q + j], 0);
                });
            }
        });
    }
}

// TESTSORTED: is the loaded slot already non-decreasing?  raw is in the
// striped load order (register 4 q + j of lane t = position 4 (q T + t) +
// j); positions >= valid count as the padding key.  Lane t compares its
// 16-byte groups in registers and each group's last key with the next
// position, whose key it reads from the first keys the lanes leave in the
// slot's LDS region (for t = T - 1: lane 0's next group).  Slot-uniform
// result (the lanes of a slot share one wave).
template <class E, int LOGW>
__device__ __forceinline__ bool xl_sorted(const typename E::Storage (&raw)[xl::kE], int valid, int t,
                                          typename E::Key* lk) {
    using K = typename E::Key;
    constexpr int T = 1 << (LOGW - xl::kLogE);
    constexpr int NG = xl::kE / 4;
    K k[xl::kE];
    sfor<0, xl::kE>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        k[r] = ((r / 4) * T + t) * 4 + r % 4 < valid ? E::key(raw[r]) : E::kPad;
    });
    bool ok = true;
    sfor<0, xl::kE - 1>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        if constexpr (r % 4 != 3 || T == 1) ok = ok && !(k[r + 1] < k[r]);
    });
    if constexpr (T == 1) return ok;
    constexpr int FLAG = NG * T;   // one word past the group-first keys
    sfor<0, NG>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        lk[q * T + t] = k[4 * q];
    });
    if (t == 0) lk[FLAG] = K(1);
    wave_lds_fence();
    sfor<0, NG>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        const int x = q * T + t + 1;
        const K nx = lk[x < FLAG ? x : 0];
        ok = ok && (x >= FLAG || !(nx < k[4 * q + 3]));
    });
    if (!ok) lk[FLAG] = K(0);
    wave_lds_fence();
    const bool sorted = lk[FLAG] != K(0);
    wave_lds_fence();   // the region is reused by the network
    return sorted;
}

// Sorts one array of `valid` keys (valid <= W; 0 for an idle slot) from src
// into dst with the T lanes of logical index t; lk is the array's LDS region.
// With test (TESTSORTED) a slot found sorted right after the load is stored
// unchanged without running the network (a sorted array is a fixed point of
// every layer, so the output is the same).
template <class E, int LOGW, int LDSK, bool SEG = false>
__device__ __forceinline__ void xlane_sort(const typename E::Storage* src, typename E::Storage* dst, int valid,
                                           int t, typename E::Key* lk, bool test = false) {
    typename E::Storage raw[xl::kE];
    typename E::Key out[xl::kE];
    if constexpr (SEG) xl_load_seg<E, LOGW>(raw, src, valid, t);
    else xl_load<E, LOGW>(raw, src, valid, t);
    if (test && xl_sorted<E, LOGW>(raw, valid, t, lk)) {
        sfor<0, xl::kE>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            out[r] = E::key(raw[r]);
        });
    } else {
        xl_network<E, LOGW, LDSK>(raw, out, valid, t, lk);
    }
    if constexpr (SEG) xl_store_seg<E, LOGW>(out, dst, valid, t);
    else xl_store<E, LOGW>(out, dst, valid, t);
}

// Batched launch, one wave tile (64 / T arrays) per wave: 256-thread
// workgroups, array a = in + a * n.
template <class E, int LOGW, int LDSK>
__global__ void __launch_bounds__(256) k_bitonic_xlane(const typename E::Storage* in, typename E::Storage* out,
                                                       int64_t arrays, int n, unsigned flags) {
    constexpr int LOGT = LOGW - xl::kLogE;
    constexpr int APW = 64 >> LOGT;
    constexpr int REGION = xl::region_words(LOGW);
    __shared__ __attribute__((aligned(16))) typename E::Key lds[4 * APW * REGION];
    const int tid = int(threadIdx.x);
    const int p = tid & 63;
    const int t = xl::lane_t(p, LOGT);
    const int slot = (tid >> 6) * APW + xl::lane_g(p, LOGT);
    const int64_t a = int64_t(blockIdx.x) * (4 * APW) + slot;
    const bool active = a < arrays;
    xlane_sort<E, LOGW, LDSK>(in + (active ? a : 0) * int64_t(n), out + (active ? a : 0) * int64_t(n),
                              active ? n : 0, t, lds + slot * REGION, (flags & CS_FLAG_TESTSORTED) != 0);
}

}  // namespace cs
