// Compile-time schedule of the register-tile bitonic network (see
// bitonic_tile.hpp for the design).  Plain C++17 with no HIP dependency, so
// the host-side simulator in tests/ executes exactly this plan.
#pragma once

#if defined(__HIPCC__)
#define CS_HD __host__ __device__
#else
#define CS_HD
#endif

#include "xlane_plan.hpp"

namespace cs {

// ------------------------------------------------------------------ plan --
constexpr int kMaxOps = 320;
// OP_LANESORT: every lane sorts its E slots ascending (Batcher's odd-even
// merge sort, xl::make_oem_df(E)); it stands for stages 1..e, whose only
// contract towards stage e+1 is ascending runs of E positions.
enum : int { OP_REFLECT = 0, OP_CE = 1, OP_TRANSPOSE = 2, OP_LANESORT = 3 };

struct Op {
    int kind = 0;
    int bit = 0;         // CE / REFLECT: logical bit
    int dirbit = -1;     // CE: descending where logical bit `dirbit` is set
    int cur_h = 0;       // CE / REFLECT: layout in force, L(cur_h)
    int from_h = 0;      // TRANSPOSE: source layout
    int to_h = 0;        // TRANSPOSE: target layout
    int flip_read = -1;  // TRANSPOSE: read upper half (this bit set) flipped
    int flip_write = -1; // TRANSPOSE: write upper half flipped (undo)
    int index = 0;       // TRANSPOSE: ordinal, selects the LDS buffer
};

struct Plan {
    Op ops[kMaxOps] = {};
    int nops = 0;
    int ntrans = 0;
};

constexpr Plan make_plan(int w, int e, bool lanesort = false) {
    Plan p{};
    if (e <= 0) return p;        // E = 1 has no slot bits: not a register tile
    int cur = e - 1;
    if (lanesort) {
        Op l{};
        l.kind = OP_LANESORT; l.cur_h = cur;
        p.ops[p.nops++] = l;
    }
    for (int k = lanesort ? e + 1 : 1; k <= w; ++k) {
        if (k <= e) {
            Op r{};
            r.kind = OP_REFLECT; r.bit = k - 1; r.cur_h = cur;
            p.ops[p.nops++] = r;
            for (int j = k - 2; j >= 0; --j) {
                Op c{};
                c.kind = OP_CE; c.bit = j; c.cur_h = cur;
                p.ops[p.nops++] = c;
            }
            continue;
        }
        int r = k % e;
        if (r == 0) r = e;
        int hi = k - 1;
        for (int chunk = 0; hi >= 0; ++chunk) {
            const int size = chunk == 0 ? r : e;
            const int h = chunk == 0 ? k - 1 : hi;
            Op t{};
            t.kind = OP_TRANSPOSE; t.from_h = cur; t.to_h = h;
            t.flip_read = chunk == 0 ? k - 1 : -1;
            t.flip_write = chunk == 1 ? k - 1 : -1;
            t.index = p.ntrans++;
            p.ops[p.nops++] = t;
            cur = h;
            for (int j = hi; j > hi - size; --j) {
                Op c{};
                c.kind = OP_CE; c.bit = j; c.cur_h = cur;
                c.dirbit = (chunk == 0 && j != k - 1) ? k - 1 : -1;
                p.ops[p.nops++] = c;
            }
            hi -= size;
        }
    }
    return p;
}

template <int LOGW, int LOGE, bool LANESORT = false> struct PlanOf {
    static constexpr Plan value = make_plan(LOGW, LOGE, LANESORT);
};
template <int N> struct OemOf {
    static constexpr xl::CxList value = xl::make_oem_df(N);
};

// Logical index of slot 0 of lane t in layout L(h) (e slot bits): the low
// h-e+1 lane bits keep their logical bit positions, the rest sit above h.
CS_HD constexpr int layout_base(int t, int h, int e) {
    return (t & ((1 << (h - e + 1)) - 1)) | ((t >> (h - e + 1)) << (h + 1));
}

// Transposes.  Leaving (or entering) layout L(h), slot r of lane t stands
// for logical position  transpose_base(t, ...) | transpose_slot(r, ...).
// When `flip_upper` is set the upper half (slot bit e-1, i.e. logical bit h)
// is stored with logical bits 0..h-1 inverted: the lane part of that is the
// low h-e+1 lane bits, the slot part is slot bits 0..e-2.
CS_HD constexpr bool slot_flipped(int r, int e, bool flip_upper) { return flip_upper && ((r >> (e - 1)) & 1); }
CS_HD constexpr int transpose_base(int t, int h, int e, bool flipped) {
    return layout_base(t, h, e) ^ (flipped ? ((1 << (h - e + 1)) - 1) : 0);
}
CS_HD constexpr int transpose_slot(int r, int h, int e, bool flip_upper) {
    return (slot_flipped(r, e, flip_upper) ? (r ^ ((1 << (e - 1)) - 1)) : r) << (h - e + 1);
}

// LDS word address of logical position x: one pad word every 2^padshift.
// For disjoint bit sets, lds_addr(x | y) == lds_addr(x) + lds_addr(y), so a
// lane's base and a slot's compile-time offset add up.
CS_HD constexpr int lds_addr(int x, int padshift) { return x + (x >> padshift); }

}  // namespace cs
