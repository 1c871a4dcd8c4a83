// Host execution of the cross-lane register bitonic engine
// (cuda-sort_amd/csrc/bitonic_xlane.hpp) on one 64-lane wave, with the lane
// mapping, odd-even comparator list, window layouts and LDS addresses of
// cuda-sort_amd/csrc/xlane_plan.hpp, and the cross-lane primitives as the
// hardware executes them: DPP quad_perm [1,0,3,2] / [2,3,0,1] (lane ^ 1,
// lane ^ 2), row_ror:8 (lane ^ 8 within a row of 16), v_permlane16_swap and
// v_permlane32_swap (row exchange of two registers).  It checks:
//   1. the in-lane odd-even sorter by the 0-1 principle over all 2^32 binary
//      inputs (bit-sliced; the descending sorter follows by complementing);
//   2. the merge stages: for W <= 128 every state the in-lane sorter can
//      leave for a binary input (each 32-block sorted with any count of ones),
//      which with 1. proves the network by the 0-1 principle; for every W and
//      stage k, every pair of sorted binary halves (ascending, descending) of
//      each 2^k block, so each stage is a correct merger;
//   3. random integer inputs (full range, few distinct values, INT_MIN /
//      INT_MAX) come out as std::sort of each array, through the kernel's
//      store path (LDS pass + striped read-back);
//   4. every LDS access stays inside its array's region and every read
//      returns a value written by the same pass.
// Usage: xlane_sim <logw> <ldsk> [oem|oemdf]   (run check 1 only, on the
// level-ordered or the depth-first comparator list)
#include <algorithm>
#include <array>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../cuda-sort_amd/csrc/xlane_plan.hpp"

using namespace cs::xl;

// Direction constants as the kernel holds them (the smallest / largest key);
// the simulated compare-exchange takes min or max accordingly.
enum Dir : int { kLo = 0, kHi = 1 };
static Dir flip(Dir d) { return d == kLo ? kHi : kLo; }

// Keys: int32 for integer inputs; for binary inputs a uint64 holds 64
// independent instances, one per bit (min = AND, max = OR).
static int32_t pick(int32_t a, int32_t b, Dir c) { return c == kLo ? std::min(a, b) : std::max(a, b); }
static uint64_t pick(uint64_t a, uint64_t b, Dir c) { return c == kLo ? (a & b) : (a | b); }

static bool fail(const char* what) {
    std::printf("FAIL: %s\n", what);
    return false;
}

// ------------------------------------------------------------ check 1
// Bit-sliced 0-1 test over all 2^32 inputs: bit j of lane v of word w[i] is
// the value at position i of input (hi * 256 + v * 64 + j).  Ascending only:
// the descending sorter (max to the lower position) maps a binary input x to
// the complement of what the ascending sorter makes of the complement of x,
// so it sorts every input exactly when the ascending one does.
typedef uint64_t V4 __attribute__((vector_size(32)));
static bool oem_exhaustive(const CxList& l) {
    V4 low[8];
    for (int i = 0; i < 8; ++i)
        for (int v = 0; v < 4; ++v) {
            uint64_t m = 0;
            for (int j = 0; j < 64; ++j) m |= uint64_t(((v * 64 + j) >> i) & 1) << j;
            low[i][v] = m;
        }
    for (uint64_t hi = 0; hi < (uint64_t(1) << 24); ++hi) {
        V4 w[32];
        for (int i = 0; i < 8; ++i) w[i] = low[i];
        for (int i = 8; i < 32; ++i) {
            const uint64_t b = ((hi >> (i - 8)) & 1) ? ~uint64_t(0) : 0;
            w[i] = V4{b, b, b, b};
        }
        for (int c = 0; c < l.n; ++c) {
            const V4 a = w[l.c[c].a], b = w[l.c[c].b];
            w[l.c[c].a] = a & b;
            w[l.c[c].b] = a | b;
        }
        V4 bad = V4{0, 0, 0, 0};
        for (int i = 0; i + 1 < 32; ++i) bad |= w[i] & ~w[i + 1];   // a 1 before a 0
        if (bad[0] | bad[1] | bad[2] | bad[3]) return fail("odd-even sorter leaves a binary input unsorted");
    }
    return true;
}

// ------------------------------------------------------------ the wave
template <class Key>
struct Wave {
    int logw, ldsk, logt, T, apw, region;
    std::vector<std::array<Key, kE>> k;   // [physical lane][register]
    std::vector<Key> lds;
    std::vector<int> stamp;               // pass number that wrote each LDS word
    int pass = 0;

    Wave(int w, int d) : logw(w), ldsk(d), logt(w - kLogE), T(1 << (w - kLogE)), apw(64 >> (w - kLogE)),
                         region(region_words(w)), k(64), lds(size_t(64 >> (w - kLogE)) * region_words(w)),
                         stamp(lds.size(), -1) {}

    int t_of(int p) const { return lane_t(p, logt); }
    int g_of(int p) const { return lane_g(p, logt); }

    bool lds_put(int p, int addr, Key v) {
        if (addr < 0 || addr >= region) return fail("LDS write outside the array region");
        const size_t a = size_t(g_of(p)) * region + addr;
        lds[a] = v;
        stamp[a] = pass;
        return true;
    }
    bool lds_get(int p, int addr, Key& v) {
        if (addr < 0 || addr >= region) return fail("LDS read outside the array region");
        const size_t a = size_t(g_of(p)) * region + addr;
        if (stamp[a] != pass) return fail("LDS read of a word this pass did not write");
        v = lds[a];
        return true;
    }

    // direction constant of lane p in stage ks (kLo = ascending choice)
    Dir cl_of(int p, int ks) const {
        const bool last = ks == logw;
        return (!last && ((t_of(p) >> (ks - 5)) & 1)) ? kHi : kLo;
    }

    void lane_sort() {
        for (int p = 0; p < 64; ++p) {
            const Dir cl = logw == kLogE ? kLo : ((t_of(p) & 1) ? kHi : kLo);
            for (int c = 0; c < kOem32.n; ++c) {
                Key& a = k[p][kOem32.c[c].a];
                Key& b = k[p][kOem32.c[c].b];
                const Key x = a, y = b;
                a = pick(x, y, cl);
                b = pick(x, y, flip(cl));
            }
        }
    }
    void reg_layer(int j, int ks) {
        for (int p = 0; p < 64; ++p) {
            const Dir cl = cl_of(p, ks);
            for (int r = 0; r < kE; ++r)
                if (((r >> j) & 1) == 0) {
                    const Key x = k[p][r], y = k[p][r | (1 << j)];
                    k[p][r] = pick(x, y, cl);
                    k[p][r | (1 << j)] = pick(x, y, flip(cl));
                }
        }
    }
    static int dpp_src(int pos, int p) {
        if (pos == 0) return (p & ~3) | ((p & 3) ^ 1);         // quad_perm [1,0,3,2]
        if (pos == 1) return (p & ~3) | ((p & 3) ^ 2);         // quad_perm [2,3,0,1]
        return (p & ~15) | ((p + 8) & 15);                     // row_ror:8
    }
    // permlane16/32_swap(a, b) as measured on the GPU: returns (first, second)
    static void rows_swap(int pos, const std::vector<Key>& a, const std::vector<Key>& b, std::vector<Key>& f,
                          std::vector<Key>& s) {
        const int d = pos == 4 ? 16 : 32;
        for (int p = 0; p < 64; ++p) {
            const bool upper = (p & d) != 0;
            f[p] = upper ? b[p - d] : a[p];
            s[p] = upper ? b[p] : a[p + d];
        }
    }
    void lane_layer(int L, int ks) {
        const int pos = kPos[L];
        if (dpp_pos(pos)) {
            std::vector<std::array<Key, kE>> nk = k;
            for (int p = 0; p < 64; ++p) {
                const Dir cl = cl_of(p, ks);
                const Dir c = ((t_of(p) >> L) & 1) ? flip(cl) : cl;
                for (int r = 0; r < kE; ++r) nk[p][r] = pick(k[p][r], k[dpp_src(pos, p)][r], c);
            }
            k = nk;
        } else {
            std::vector<Key> a(64), b(64), f(64), s(64), lo(64), hi(64);
            for (int r = 0; r < kE; r += 2) {
                for (int p = 0; p < 64; ++p) { a[p] = k[p][r]; b[p] = k[p][r + 1]; }
                rows_swap(pos, a, b, f, s);
                for (int p = 0; p < 64; ++p) {
                    const Dir cl = cl_of(p, ks);
                    lo[p] = pick(f[p], s[p], cl);
                    hi[p] = pick(f[p], s[p], flip(cl));
                }
                rows_swap(pos, lo, hi, f, s);
                for (int p = 0; p < 64; ++p) { k[p][r] = f[p]; k[p][r + 1] = s[p]; }
            }
        }
    }
    bool put_std() {
        ++pass;
        for (int p = 0; p < 64; ++p)
            for (int r = 0; r < kE; ++r)
                if (!lds_put(p, lds_addr(t_of(p) << kLogE) + lds_addr(r), k[p][r])) return false;
        return true;
    }
    bool get_std() {
        for (int p = 0; p < 64; ++p)
            for (int r = 0; r < kE; ++r)
                if (!lds_get(p, lds_addr(t_of(p) << kLogE) + lds_addr(r), k[p][r])) return false;
        return true;
    }
    bool put_win(int ks) {
        ++pass;
        for (int p = 0; p < 64; ++p)
            for (int s = 0; s < kE; ++s)
                if (!lds_put(p, lds_addr(win_lane_pos(t_of(p), ks)) + lds_addr(win_slot_pos(s, ks)), k[p][s]))
                    return false;
        return true;
    }
    bool get_win(int ks) {
        for (int p = 0; p < 64; ++p)
            for (int s = 0; s < kE; ++s)
                if (!lds_get(p, lds_addr(win_lane_pos(t_of(p), ks)) + lds_addr(win_slot_pos(s, ks)), k[p][s]))
                    return false;
        return true;
    }
    bool stage(int ks) {
        if (ks < ldsk) {
            for (int L = ks - 6; L >= 0; --L) lane_layer(L, ks);
            for (int j = 4; j >= 0; --j) reg_layer(j, ks);
        } else {
            if (!put_std() || !get_win(ks)) return false;
            for (int j = 4; j >= 0; --j) reg_layer(j, ks);
            if (!put_win(ks) || !get_std()) return false;
            for (int L = ks - 11; L >= 0; --L) lane_layer(L, ks);
            for (int j = std::min(ks - 6, 4); j >= 0; --j) reg_layer(j, ks);
        }
        return true;
    }
    // Load: group q of lane t is positions 4 (q T + t) .. + 3, declared as
    // registers 4q .. 4q+3 of the standard layout (as the kernel does).
    void load(const std::vector<std::vector<Key>>& arrays) {
        for (int p = 0; p < 64; ++p)
            for (int q = 0; q < kE / 4; ++q)
                for (int j = 0; j < 4; ++j) k[p][4 * q + j] = arrays[g_of(p)][(q * T + t_of(p)) * 4 + j];
    }
    // Store path: standard layout into LDS, group q read back striped.
    bool store(std::vector<std::vector<Key>>& out) {
        if (!put_std()) return false;
        out.assign(apw, std::vector<Key>(size_t(1) << logw));
        for (int p = 0; p < 64; ++p)
            for (int q = 0; q < kE / 4; ++q) {
                const int c = q * T + t_of(p);
                for (int j = 0; j < 4; ++j) {
                    Key v;
                    if (!lds_get(p, lds_addr(4 * c) + j, v)) return false;
                    out[g_of(p)][4 * c + j] = v;
                }
            }
        return true;
    }
    bool sort(const std::vector<std::vector<Key>>& in, std::vector<std::vector<Key>>& out) {
        load(in);
        lane_sort();
        for (int ks = kLogE + 1; ks <= logw; ++ks)
            if (!stage(ks)) return false;
        return store(out);
    }
};

static bool check_lane_map(int logt) {
    std::vector<int> seen(64, 0);
    for (int p = 0; p < 64; ++p) {
        if (lane_of(lane_t(p, logt), lane_g(p, logt), logt) != p) return fail("lane_of is not the inverse of (lane_t, lane_g)");
        seen[lane_t(p, logt) * (64 >> logt) + lane_g(p, logt)]++;
    }
    for (int v : seen)
        if (v != 1) return fail("lane mapping is not a bijection");
    return true;
}

// ------------------------------------------------------------ check 2
// Stage ks alone on a tile whose 2^ks blocks hold a sorted binary ascending
// half and a sorted binary descending half (ones counts ca, cb), i.e. every
// input a correct previous stage can produce.  The block direction follows
// logical bit ks.  Registers are placed in the standard layout directly.
static bool stage_merger(int logw, int ldsk, int ks) {
    using Key = uint64_t;
    Wave<Key> wv(logw, ldsk);
    const int half = 1 << (ks - 1), W = 1 << logw;
    const int blocks = (W >> ks) * wv.apw * 64;   // block instances per run
    const int combos = (half + 1) * (half + 1);
    std::vector<std::vector<Key>> lg(wv.apw, std::vector<Key>(W));   // logical images
    for (int base = 0; base < combos; base += blocks) {
        for (int a = 0; a < wv.apw; ++a)
            for (int b = 0; b < (W >> ks); ++b) {
                for (int i = 0; i < 2 * half; ++i) lg[a][b * 2 * half + i] = 0;
                for (int bit = 0; bit < 64; ++bit) {
                    const int combo = (base + (a * (W >> ks) + b) * 64 + bit) % combos;
                    const int ca = combo / (half + 1), cb = combo % (half + 1);
                    // first half ascending (ones at the end), second descending
                    for (int i = 0; i < half; ++i) {
                        if (i >= half - ca) lg[a][b * 2 * half + i] |= Key(1) << bit;
                        if (i < cb) lg[a][b * 2 * half + half + i] |= Key(1) << bit;
                    }
                }
            }
        for (int p = 0; p < 64; ++p)
            for (int r = 0; r < kE; ++r) wv.k[p][r] = lg[wv.g_of(p)][(wv.t_of(p) << kLogE) + r];
        if (!wv.stage(ks)) return false;
        for (int p = 0; p < 64; ++p)
            for (int r = 0; r < kE; ++r) lg[wv.g_of(p)][(wv.t_of(p) << kLogE) + r] = wv.k[p][r];
        for (int a = 0; a < wv.apw; ++a)
            for (int b = 0; b < (W >> ks); ++b) {
                const bool desc = ks < logw && ((b >> 0) & 1);   // logical bit ks of the block start
                const Key* v = &lg[a][b * 2 * half];
                for (int i = 0; i + 1 < 2 * half; ++i)
                    if (desc ? (~v[i] & v[i + 1]) : (v[i] & ~v[i + 1])) {
                        std::printf("stage %d, block %d: ", ks, b);
                        return fail("merge stage leaves a block unsorted");
                    }
            }
    }
    return true;
}

// Every binary state the in-lane sorter can leave (W <= 128): each 32-block
// sorted in its direction with any number of ones, through all stages.
static bool binary_states(int logw, int ldsk) {
    using Key = int32_t;
    Wave<Key> wv(logw, ldsk);
    const int W = 1 << logw, nb = W / kE;
    long states = 1;
    for (int b = 0; b < nb; ++b) states *= kE + 1;
    std::vector<std::vector<Key>> in(wv.apw, std::vector<Key>(W)), out;
    for (long s0 = 0; s0 < states; s0 += wv.apw) {
        for (int a = 0; a < wv.apw; ++a) {
            long s = (s0 + a) % states;
            for (int b = 0; b < nb; ++b) {
                const int ones = int(s % (kE + 1));
                s /= kE + 1;
                for (int i = 0; i < kE; ++i) in[a][b * kE + i] = i >= kE - ones ? 1 : 0;
            }
        }
        // place as the state after lane_sort: block b is lane t = b (standard
        // layout), ascending where lane bit 0 is 0, descending otherwise
        for (int p = 0; p < 64; ++p) {
            const int t = wv.t_of(p);
            for (int r = 0; r < kE; ++r) {
                const Key v = in[wv.g_of(p)][t * kE + r];
                wv.k[p][(nb > 1 && (t & 1)) ? kE - 1 - r : r] = v;
            }
        }
        for (int ks = kLogE + 1; ks <= logw; ++ks)
            if (!wv.stage(ks)) return false;
        if (!wv.store(out)) return false;
        for (int a = 0; a < wv.apw; ++a) {
            int ones = 0;
            for (Key v : in[a]) ones += v;
            for (int i = 0; i < W; ++i)
                if (out[a][i] != (i >= W - ones ? 1 : 0)) return fail("binary state not sorted");
        }
    }
    return true;
}

// ------------------------------------------------------------ check 3
static bool random_inputs(int logw, int ldsk, int trials) {
    using Key = int32_t;
    constexpr Key kMin = INT32_MIN, kMax = INT32_MAX;
    Wave<Key> wv(logw, ldsk);
    const int W = 1 << logw;
    std::mt19937_64 rng(1234 + logw * 100 + ldsk);
    std::vector<std::vector<Key>> in(wv.apw, std::vector<Key>(W)), out;
    for (int tr = 0; tr < trials; ++tr) {
        const int kind = tr % 4;
        for (auto& a : in)
            for (auto& v : a) {
                const uint64_t x = rng();
                if (kind == 0) v = Key(uint32_t(x));
                else if (kind == 1) v = Key(x % 5) - 2;
                else if (kind == 2) v = (x & 1) ? kMax : kMin;
                else v = Key(uint32_t(x >> 7) % 1000u);
            }
        if (!wv.sort(in, out)) return false;
        for (int a = 0; a < wv.apw; ++a) {
            std::vector<Key> ref = in[a];
            std::sort(ref.begin(), ref.end());
            if (ref != out[a]) return fail("random input not sorted to std::sort's sequence");
        }
    }
    return true;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: xlane_sim <logw> <ldsk> [oem]\n");
        return 2;
    }
    const int logw = std::atoi(argv[1]), ldsk = std::atoi(argv[2]);
    // "oem": the engine's level-ordered list; "oemdf": the depth-first order
    // of the register tile's in-lane opening (make_oem_df)
    if (argc > 3) return oem_exhaustive(std::string(argv[3]) == "oemdf" ? make_oem_df(32) : kOem32) ? 0 : 1;
    if (logw < 5 || logw > 11) return fail("logw outside 5..11") ? 0 : 1;
    if (logw == 11 && ldsk > 11) return fail("2048 positions need the window for stage 11") ? 0 : 1;
    if (!check_lane_map(logw - kLogE)) return 1;
    for (int ks = kLogE + 1; ks <= logw; ++ks)
        if (!stage_merger(logw, ldsk, ks)) return 1;
    if (logw <= 7 && !binary_states(logw, ldsk)) return 1;
    if (!random_inputs(logw, ldsk, logw >= 9 ? 24 : 64)) return 1;
    std::printf("ok logw=%d ldsk=%d\n", logw, ldsk);
    return 0;
}
