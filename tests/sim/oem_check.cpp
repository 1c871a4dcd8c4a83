// 0-1 principle check of the odd-even merge presort network (plan.hpp):
// every 0/1 input of n wires comes out sorted, bit-sliced 64 inputs at a
// time (exhaustive: 2^n inputs).  Also checks the comparator counts.
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../../cuda-sort_amd/csrc/plan.hpp"

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 16;
    cs::CePairs p = cs::make_oem(n);
    const int drop = argc > 2 ? std::atoi(argv[2]) : -1;   // planted mistake: remove one comparator
    if (drop >= 0 && drop < p.n) {
        for (int c = drop; c + 1 < p.n; ++c) {
            p.a[c] = p.a[c + 1];
            p.b[c] = p.b[c + 1];
        }
        --p.n;
    }
    const int want = (n == 2 ? 1 : n == 4 ? 5 : n == 8 ? 19 : n == 16 ? 63 : n == 32 ? 191 : -1) - (drop >= 0);
    if (p.n != want) {
        std::printf("comparators %d, expected %d\n", p.n, want);
        return 1;
    }
    // inputs x = hi * 64 + lane: wire j of lane l holds bit j of x
    const uint64_t total = n >= 64 ? 0 : (uint64_t(1) << n);
    const uint64_t batches = total < 64 ? 1 : total / 64;
    for (uint64_t hi = 0; hi < batches; ++hi) {
        uint64_t w[32];
        for (int j = 0; j < n; ++j) {
            if (j < 6) {
                uint64_t m = 0;
                for (int l = 0; l < 64; ++l)
                    if ((l >> j) & 1) m |= uint64_t(1) << l;
                w[j] = m;
            } else {
                w[j] = ((hi >> (j - 6)) & 1) ? ~uint64_t(0) : 0;
            }
        }
        for (int c = 0; c < p.n; ++c) {
            const uint64_t x = w[p.a[c]], y = w[p.b[c]];
            w[p.a[c]] = x & y;   // min
            w[p.b[c]] = x | y;   // max
        }
        const uint64_t valid = total < 64 ? ((uint64_t(1) << total) - 1) : ~uint64_t(0);
        for (int j = 0; j + 1 < n; ++j) {
            if ((w[j] & ~w[j + 1]) & valid) {   // a 1 above a 0
                std::printf("unsorted output for batch %llu at wire %d\n", (unsigned long long)hi, j);
                return 1;
            }
        }
    }
    std::printf("ok %d wires, %d comparators\n", n, p.n);
    return 0;
}
