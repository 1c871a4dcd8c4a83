// Randomised check of the serving pool's offset allocator (cuda-sort_amd/csrc/pool.hpp)
// against a byte-map model: no two live blocks overlap, every block lies inside
// the capacity and is aligned, in_use/peak accounting is exact, a request
// larger than the largest free block fails, and freeing everything coalesces
// back to a single block of the full capacity.
#include <cstdio>
#include <map>
#include <random>
#include <vector>

#include "../../cuda-sort_amd/csrc/pool.hpp"

int main() {
    std::mt19937_64 rng(7);
    for (int round = 0; round < 50; ++round) {
        const size_t cap = (1 + rng() % 64) * 4096 + (rng() % 255);
        cs::OffsetPool pool(cap);
        const size_t real_cap = cap / cs::OffsetPool::kAlign * cs::OffsetPool::kAlign;
        std::vector<char> owner(real_cap, 0);
        std::map<size_t, size_t> live;   // offset -> rounded size
        size_t in_use = 0, peak = 0;
        for (int step = 0; step < 2000; ++step) {
            if (live.empty() || rng() % 3) {
                const size_t want = rng() % (real_cap / 4 + 1);
                size_t off = 0;
                const size_t largest = pool.largest_free();
                const bool ok = pool.alloc(want, &off);
                const size_t need = cs::OffsetPool::round_up(want ? want : 1);
                if (ok != (need <= largest)) { std::printf("alloc success does not match largest free block\n"); return 1; }
                if (!ok) continue;
                if (off % cs::OffsetPool::kAlign || off + need > real_cap) { std::printf("bad block\n"); return 1; }
                for (size_t b = off; b < off + need; ++b) {
                    if (owner[b]) { std::printf("overlap\n"); return 1; }
                    owner[b] = 1;
                }
                live[off] = need;
                in_use += need;
                peak = in_use > peak ? in_use : peak;
            } else {
                auto it = live.begin();
                std::advance(it, long(rng() % live.size()));
                for (size_t b = it->first; b < it->first + it->second; ++b) owner[b] = 0;
                if (!pool.release(it->first)) { std::printf("release failed\n"); return 1; }
                in_use -= it->second;
                live.erase(it);
            }
            if (pool.in_use() != in_use || pool.peak() != peak) { std::printf("accounting\n"); return 1; }
        }
        if (pool.release(real_cap + 1)) { std::printf("unknown offset released\n"); return 1; }
        for (auto& kv : live) pool.release(kv.first);
        if (pool.in_use() != 0 || pool.free_blocks() != (real_cap ? 1u : 0u) || pool.largest_free() != real_cap) {
            std::printf("not coalesced\n");
            return 1;
        }
    }
    std::printf("ok\n");
    return 0;
}
