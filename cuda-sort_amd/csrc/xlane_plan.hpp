// Schedule of the cross-lane register bitonic engine (bitonic_xlane.hpp) for
// 4-byte keys without payload: the lane mapping, the in-lane odd-even merge
// sorter, the LDS window layouts and their addresses.  Plain C++17 with no HIP
// dependency, so the host simulator in tests/sim/ executes the same helpers.
#pragma once

#if defined(__HIPCC__)
#define CS_XL_HD __host__ __device__
#else
#define CS_XL_HD
#endif

namespace cs {
namespace xl {

// 32 keys per lane: logical bits 0..4 of a position are register bits.
constexpr int kLogE = 5;
constexpr int kE = 1 << kLogE;

// Physical lane position (bit of the lane id within the wave) that carries
// logical bit 5 + i.  Positions 0, 1 and 3 are a single DPP pattern away
// (lane ^ 1, ^ 2: quad_perm; lane ^ 8: row_ror:8), 4 and 5 are crossed with
// v_permlane16_swap / v_permlane32_swap; position 2 (lane ^ 4, no single DPP
// pattern) is used last, so that for up to 32 lanes per array it carries the
// array slot instead of a key bit.
constexpr int kPos[6] = {0, 1, 3, 4, 5, 2};

CS_XL_HD constexpr bool dpp_pos(int pos) { return pos == 0 || pos == 1 || pos == 3; }

// Logical lane index (logt bits) of physical lane p.
CS_XL_HD constexpr int lane_t(int p, int logt) {
    int t = 0;
    for (int i = 0; i < logt; ++i) t |= ((p >> kPos[i]) & 1) << i;
    return t;
}
// Array slot of physical lane p: the lane-id bits not used by lane_t, in
// increasing position order.
CS_XL_HD constexpr int lane_g(int p, int logt) {
    int g = 0, o = 0;
    for (int pos = 0; pos < 6; ++pos) {
        bool used = false;
        for (int i = 0; i < logt; ++i) used = used || kPos[i] == pos;
        if (!used) g |= ((p >> pos) & 1) << o++;
    }
    return g;
}
// Physical lane of (t, g): the inverse of (lane_t, lane_g).
CS_XL_HD constexpr int lane_of(int t, int g, int logt) {
    int p = 0, o = 0;
    for (int i = 0; i < logt; ++i) p |= ((t >> i) & 1) << kPos[i];
    for (int pos = 0; pos < 6; ++pos) {
        bool used = false;
        for (int i = 0; i < logt; ++i) used = used || kPos[i] == pos;
        if (!used) p |= ((g >> o++) & 1) << pos;
    }
    return p;
}

// LDS word of logical position x inside an array's region: 4 pad words per
// 32 keep every lane's 32 consecutive keys 16-byte aligned (ds_write_b128 /
// ds_read_b128) and 4 banks apart from its neighbour's.  Additive over
// disjoint bit sets: lds_addr(a | b) == lds_addr(a) + lds_addr(b).
CS_XL_HD constexpr int lds_addr(int x) { return x + 4 * (x >> 5); }
// Words per array region.  Eight consecutive lanes (one 128-byte LDS cycle of
// 16-byte accesses) hold min(T, 4) lanes of each of 8 / min(T, 4) arrays; the
// lanes of one array are 36 words (4 banks) apart, so consecutive regions
// start 4 * min(T, 4) banks apart (mod 32) and the eight accesses fall on
// distinct bank groups.  The smallest size >= the padded array with that
// residue.
constexpr int region_words(int logw) {
    const int t = 1 << (logw - kLogE);
    const int off = 4 * (t < 4 ? t : 4);
    int r = lds_addr(1 << logw);
    while (r % 32 != off % 32) ++r;
    return r;
}

// Standard layout: lane t holds logical positions 32 t + r in register r.
// Window layout of stage k (k >= 6): register s holds logical bit k - 5 + s
// (the five bits k-5 .. k-1); the lane index u holds the other logical bits,
// low ones first: bits 0 .. k-6, then bits k .. logw-1.
CS_XL_HD constexpr int win_lane_pos(int u, int k) {
    const int low = k - 5;
    return (u & ((1 << low) - 1)) | ((u >> low) << k);
}
CS_XL_HD constexpr int win_slot_pos(int s, int k) { return s << (k - 5); }

// Batcher's odd-even merge sort on 32 inputs: 191 comparators (the bitonic
// sorter needs 240).  Comparator (a, b), a < b, puts the smaller key at a.
struct Cx { int a = 0, b = 0; };
struct CxList {
    Cx c[200] = {};
    int n = 0;
};
constexpr CxList make_oem(int n) {
    CxList l{};
    for (int p = 1; p < n; p <<= 1)
        for (int k = p; k >= 1; k >>= 1)
            for (int j = k % p; j + k < n; j += 2 * k)
                for (int i = 0; i < k && i + j + k < n; ++i)
                    if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                        l.c[l.n].a = i + j;
                        l.c[l.n].b = i + j + k;
                        ++l.n;
                    }
    return l;
}
// The same comparators listed depth first (sort the lower half, sort the
// upper half, merge): the register tile's in-lane opening uses this order
// (measured on IntAndData N = 129: 38.56 -> 38.18 us); the cross-lane
// engine keeps the level order above (its 4096 x 1024 batch read 12.04 ->
// 12.07 us with this one).
constexpr void oem_merge(CxList& l, int lo, int n, int r) {
    const int step = r * 2;
    if (step < n) {
        oem_merge(l, lo, n, step);
        oem_merge(l, lo + r, n, step);
        for (int i = lo + r; i + r < lo + n; i += step) {
            l.c[l.n].a = i;
            l.c[l.n].b = i + r;
            ++l.n;
        }
    } else {
        l.c[l.n].a = lo;
        l.c[l.n].b = lo + r;
        ++l.n;
    }
}
constexpr void oem_sort(CxList& l, int lo, int n) {
    if (n > 1) {
        oem_sort(l, lo, n / 2);
        oem_sort(l, lo + n / 2, n / 2);
        oem_merge(l, lo, n, 1);
    }
}
constexpr CxList make_oem_df(int n) {
    CxList l{};
    oem_sort(l, 0, n);
    return l;
}
constexpr CxList kOem32 = make_oem(32);
static_assert(kOem32.n == 191, "odd-even merge sort of 32 has 191 comparators");

}  // namespace xl
}  // namespace cs
