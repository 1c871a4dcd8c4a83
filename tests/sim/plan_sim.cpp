// Host execution of the register-tile bitonic schedule (cuda-sort_amd/csrc/plan.hpp),
// lane by lane and slot by slot, with the same layout, flip and LDS-address
// helpers the kernel uses, with or without the in-lane odd-even opening
// (third argument 1: OP_LANESORT, the form records and 8-byte keys run).
// It checks, for a (w, e) tiling:
//   1. the 0-1 principle: every one of the 2^W binary inputs comes out
//      sorted (W <= 16), else many random binary inputs;
//   2. random integer inputs with many duplicates come out sorted and are a
//      permutation of the input;
//   3. records: payloads survive (each exactly once), and max-key padding at
//      positions >= n never enters the valid prefix even when valid keys
//      equal the padding key (the reference's tie/padding guarantee);
//   4. lds_addr is additive over the lane/slot split used by the kernel.
// Exit status 0 on success; a message names the first failure.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../cuda-sort_amd/csrc/plan.hpp"

using namespace cs;

struct Sim {
    int w, e, T, E, W;
    Plan plan;
    std::vector<int64_t> k, p;   // [t * E + r]
    std::vector<int64_t> lk, lp;
    Sim(int w_, int e_, bool lanesort)
        : w(w_), e(e_), T(1 << (w_ - e_)), E(1 << e_), W(1 << w_), plan(make_plan(w_, e_, lanesort)) {
        lk.resize(size_t(W) * 2 + 64);
        lp.resize(lk.size());
    }
    int64_t& K(int t, int r) { return k[size_t(t) * E + r]; }
    int64_t& P(int t, int r) { return p[size_t(t) * E + r]; }
    void cx(int t, int a, int b, bool desc) {
        int64_t& ka = K(t, a); int64_t& kb = K(t, b);
        const bool swap = desc ? (ka < kb) : (kb < ka);
        if (swap) { std::swap(ka, kb); std::swap(P(t, a), P(t, b)); }
    }
    bool run(const std::vector<int64_t>& keys, const std::vector<int64_t>& pay) {
        k.assign(keys.begin(), keys.end());   // identity layout L(e-1)
        p.assign(pay.begin(), pay.end());
        for (int i = 0; i < plan.nops; ++i) {
            const Op& op = plan.ops[i];
            if (op.kind == OP_LANESORT) {
                const xl::CxList l = xl::make_oem_df(E);
                for (int t = 0; t < T; ++t)
                    for (int c = 0; c < l.n; ++c) cx(t, l.c[c].a, l.c[c].b, false);
            } else if (op.kind == OP_REFLECT) {
                const int mask = (2 << op.bit) - 1;
                for (int t = 0; t < T; ++t)
                    for (int r = 0; r < E; ++r)
                        if (((r >> op.bit) & 1) == 0) cx(t, r, r ^ mask, false);
            } else if (op.kind == OP_CE) {
                const int s = op.bit - (op.cur_h - e + 1);
                if (s < 0 || s >= e) { std::printf("compare on a lane bit\n"); return false; }
                const int sd = op.dirbit >= 0 ? op.dirbit - (op.cur_h - e + 1) : -1;
                if (op.dirbit >= 0 && (sd < 0 || sd >= e)) { std::printf("direction on a lane bit\n"); return false; }
                for (int t = 0; t < T; ++t)
                    for (int r = 0; r < E; ++r)
                        if (((r >> s) & 1) == 0) cx(t, r, r | (1 << s), sd >= 0 && ((r >> sd) & 1));
            } else {
                const int PS = 5;
                std::fill(lk.begin(), lk.end(), -999);
                std::vector<char> used(lk.size(), 0);
                for (int t = 0; t < T; ++t)
                    for (int r = 0; r < E; ++r) {
                        const bool f = slot_flipped(r, e, op.flip_write >= 0);
                        const int tb = transpose_base(t, op.from_h, e, f);
                        const int sl = transpose_slot(r, op.from_h, e, op.flip_write >= 0);
                        if (tb & sl) { std::printf("lane and slot parts overlap\n"); return false; }
                        const int a = lds_addr(tb, PS) + lds_addr(sl, PS);
                        if (a != lds_addr(tb | sl, PS)) { std::printf("lds_addr not additive\n"); return false; }
                        if (used[size_t(a)]) { std::printf("two slots write one LDS word\n"); return false; }
                        used[size_t(a)] = 1;
                        lk[size_t(a)] = K(t, r);
                        lp[size_t(a)] = P(t, r);
                    }
                for (int t = 0; t < T; ++t)
                    for (int r = 0; r < E; ++r) {
                        const bool f = slot_flipped(r, e, op.flip_read >= 0);
                        const int a = lds_addr(transpose_base(t, op.to_h, e, f), PS) +
                                      lds_addr(transpose_slot(r, op.to_h, e, op.flip_read >= 0), PS);
                        if (!used[size_t(a)]) { std::printf("read of an unwritten LDS word\n"); return false; }
                        K(t, r) = lk[size_t(a)];
                        P(t, r) = lp[size_t(a)];
                    }
            }
        }
        return true;   // output: logical t*E + r, the identity layout
    }
};

int main(int argc, char** argv) {
    if (argc < 3) { std::printf("usage: plan_sim w e [lanesort]\n"); return 2; }
    const int w = std::atoi(argv[1]), e = std::atoi(argv[2]);
    Sim sim(w, e, argc > 3 && std::atoi(argv[3]) != 0);
    const int W = sim.W;
    std::vector<int64_t> keys(static_cast<size_t>(W)), pay(static_cast<size_t>(W));
    // 1. zero-one principle
    std::mt19937_64 rng(12345);
    const bool exhaustive = W <= 16;
    // work per case ~ W * ops; keep each tiling to a fraction of a second
    const uint64_t budget = uint64_t(1) << 24;
    const uint64_t per = uint64_t(W) * uint64_t(sim.plan.nops + 1);
    const uint64_t cases = exhaustive ? (uint64_t(1) << W) : std::max<uint64_t>(16, std::min<uint64_t>(20000, budget / per));
    const int trials = int(std::max<uint64_t>(8, std::min<uint64_t>(300, budget / per)));
    for (uint64_t c = 0; c < cases; ++c) {
        const uint64_t bits = exhaustive ? c : rng();
        for (int i = 0; i < W; ++i) {
            keys[size_t(i)] = exhaustive ? int64_t((bits >> i) & 1) : int64_t(rng() & 1);
            pay[size_t(i)] = i;
        }
        if (!sim.run(keys, pay)) return 1;
        for (int i = 1; i < W; ++i)
            if (sim.k[size_t(i)] < sim.k[size_t(i) - 1]) {
                std::printf("w=%d e=%d: 0-1 input %llu not sorted at %d\n", w, e, (unsigned long long)c, i);
                return 1;
            }
    }
    // 2./3. random keys with duplicates and padding
    for (int trial = 0; trial < trials; ++trial) {
        const int n = trial == 0 ? W : int(1 + rng() % uint64_t(W));
        const int64_t range = 1 + int64_t(rng() % 8);
        const int64_t pad = 1000;
        for (int i = 0; i < W; ++i) {
            if (i < n) {
                // valid keys include the padding value itself
                keys[size_t(i)] = (rng() % 5 == 0) ? pad : int64_t(rng() % uint64_t(range));
                pay[size_t(i)] = i;
            } else {
                keys[size_t(i)] = pad;
                pay[size_t(i)] = -1;
            }
        }
        if (!sim.run(keys, pay)) return 1;
        std::vector<int64_t> want(keys.begin(), keys.begin() + n);
        std::sort(want.begin(), want.end());
        std::vector<int> seen(static_cast<size_t>(n), 0);
        for (int i = 0; i < n; ++i) {
            if (sim.k[size_t(i)] != want[size_t(i)]) { std::printf("w=%d e=%d: key %d wrong\n", w, e, i); return 1; }
            const int64_t o = sim.p[size_t(i)];
            if (o < 0) { std::printf("w=%d e=%d: padding entered the valid prefix at %d\n", w, e, i); return 1; }
            if (keys[size_t(o)] != sim.k[size_t(i)]) { std::printf("w=%d e=%d: payload/key mismatch\n", w, e); return 1; }
            if (seen[size_t(o)]++) { std::printf("w=%d e=%d: payload duplicated\n", w, e); return 1; }
        }
    }
    std::printf("ok w=%d e=%d ops=%d transposes=%d\n", w, e, sim.plan.nops, sim.plan.ntrans);
    return 0;
}
