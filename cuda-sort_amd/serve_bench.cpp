// Open-loop load generator for the serving layer (cs_server_*).
//
// Requests arrive as a Poisson process of rate --rate per second for
// --seconds; request sizes are floor(exp(U[ln min, ln(max+1)))) clipped to
// [min, max]; elements are records {int64 key, int32 payload = index within
// the request} with keys uniform over 64 bits.  The main thread only
// submits: each request at its scheduled arrival time, with that time as its
// arrival stamp, so any lateness counts as latency.  Buffers come from a
// ring of --ring slots; --workers threads own the ring slots (slot k belongs
// to worker k mod workers) and, for each request of their slots in turn,
// wait for it, check every --verify-every-th result (sorted keys, payload
// permutation, key/payload association), release it and fill the slot with
// the input of the request that reuses it.  Prints one JSON line.
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "../include/cusort_hip.h"

namespace {

uint64_t mix(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t x) { return double(mix(x) >> 11) * 0x1p-53; }

struct Rec { int64_t key; int32_t payload; int32_t pad; };

bool verify(const Rec* in, const Rec* out, int64_t n) {
    std::vector<char> seen(static_cast<size_t>(n), 0);
    for (int64_t i = 0; i < n; ++i) {
        if (i && out[i].key < out[i - 1].key) return false;
        const int64_t p = out[i].payload;
        if (p < 0 || p >= n || seen[size_t(p)]++) return false;
        if (in[p].key != out[i].key) return false;
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    double rate = 20000, seconds = 10, budget_gb = 32;
    int64_t min_n = 16, max_n = 65536;
    int slots = 4, ring = 1024, verify_every = 1, workers = 8;
    uint64_t seed = 42;
    for (int i = 1; i + 1 < argc; i += 2) {
        const std::string k = argv[i];
        const char* v = argv[i + 1];
        if (k == "--rate") rate = std::atof(v);
        else if (k == "--seconds") seconds = std::atof(v);
        else if (k == "--budget-gb") budget_gb = std::atof(v);
        else if (k == "--min") min_n = std::atoll(v);
        else if (k == "--max") max_n = std::atoll(v);
        else if (k == "--slots") slots = std::atoi(v);
        else if (k == "--ring") ring = std::atoi(v);
        else if (k == "--verify-every") verify_every = std::atoi(v);
        else if (k == "--workers") workers = std::atoi(v);
        else if (k == "--seed") seed = std::strtoull(v, nullptr, 10);
        else { std::fprintf(stderr, "unknown option %s\n", k.c_str()); return 2; }
    }
    if (workers < 1 || ring < workers || ring % workers) {
        std::fprintf(stderr, "--ring must be a positive multiple of --workers\n");
        return 2;
    }
    const int64_t requests = int64_t(rate * seconds);
    // arrival offsets and sizes
    std::vector<double> arrive(static_cast<size_t>(requests));
    std::vector<int64_t> size(static_cast<size_t>(requests));
    double t = 0;
    int64_t offered = 0;
    const double lo = std::log(double(min_n)), hi = std::log(double(max_n) + 1.0);
    for (int64_t r = 0; r < requests; ++r) {
        t += -std::log(1.0 - uniform01(seed ^ (uint64_t(r) * 2 + 1))) / rate;
        arrive[size_t(r)] = t;
        int64_t n = int64_t(std::floor(std::exp(lo + (hi - lo) * uniform01(seed ^ (uint64_t(r) * 2 + 2) ^ 0x5555))));
        size[size_t(r)] = n < min_n ? min_n : (n > max_n ? max_n : n);
        offered += size[size_t(r)];
    }
    cs_server* srv = nullptr;
    // 32 GB budget = 32 * 10^9 bytes (decimal)
    int rc = cs_server_create(&srv, CS_REC_I64_I32, size_t(budget_gb * 1e9), slots, 0, 0);
    if (rc != CS_OK) { std::fprintf(stderr, "cs_server_create: %s\n", cs_strerror(rc)); return 1; }
    std::vector<Rec> in(size_t(ring) * size_t(max_n)), out(size_t(ring) * size_t(max_n));
    auto fill = [&](int64_t r) {
        Rec* a = &in[size_t(r % ring) * size_t(max_n)];
        for (int64_t j = 0, n = size[size_t(r)]; j < n; ++j)
            a[j] = Rec{int64_t(mix(seed ^ (uint64_t(r) << 20) ^ uint64_t(j))), int32_t(j), 0};
    };
    // ticket[r] != 0 once submitted; ready[r] once its ring slot holds its input
    std::vector<std::atomic<uint64_t>> ticket(static_cast<size_t>(requests));
    std::vector<std::atomic<int>> ready(static_cast<size_t>(requests));
    for (int64_t r = 0; r < requests; ++r) {
        ticket[size_t(r)].store(0);
        ready[size_t(r)].store(0);
    }
    for (int64_t r = 0; r < std::min<int64_t>(ring, requests); ++r) {
        fill(r);
        ready[size_t(r)].store(1);
    }
    std::atomic<int64_t> verified{0}, failures{0};
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) {
        pool.emplace_back([&, w] {
            for (int64_t r = w; r < requests; r += workers) {
                uint64_t tk;
                while ((tk = ticket[size_t(r)].load(std::memory_order_acquire)) == 0)
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                const int st = cs_server_wait(srv, tk, -1.0);
                const size_t slot = size_t(r % ring);
                if (st != CS_OK) {
                    ++failures;
                } else if (verify_every > 0 && r % verify_every == 0) {
                    ++verified;
                    if (!verify(&in[slot * size_t(max_n)], &out[slot * size_t(max_n)], size[size_t(r)])) ++failures;
                }
                cs_server_release(srv, tk);
                if (r + ring < requests) {
                    fill(r + ring);
                    ready[size_t(r + ring)].store(1, std::memory_order_release);
                }
            }
        });
    }
    int64_t late = 0, starved = 0;
    const double t0 = cs_server_now() + 0.05;
    for (int64_t r = 0; r < requests; ++r) {
        if (!ready[size_t(r)].load(std::memory_order_acquire)) {
            ++starved;   // the ring slot is still held by an unfinished request
            while (!ready[size_t(r)].load(std::memory_order_acquire)) std::this_thread::yield();
        }
        const size_t slot = size_t(r % ring);
        const double at = t0 + arrive[size_t(r)];
        double now = cs_server_now();
        if (now > at) ++late;
        while (now < at) {
            if (at - now > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
            now = cs_server_now();
        }
        uint64_t tk = 0;
        rc = cs_server_submit(srv, &in[slot * size_t(max_n)], &out[slot * size_t(max_n)], size[size_t(r)], at, &tk);
        if (rc != CS_OK) {
            std::fprintf(stderr, "submit: %s\n", cs_strerror(rc));
            return 1;
        }
        ticket[size_t(r)].store(tk, std::memory_order_release);
    }
    for (auto& th : pool) th.join();
    cs_server_stats st{};
    cs_server_get_stats(srv, &st);
    cs_server_destroy(srv);
    std::printf("{\"requests\": %" PRId64 ", \"completed\": %" PRId64 ", \"rejected\": %" PRId64
                ", \"p50_ms\": %.4f, \"p99_ms\": %.4f, \"max_ms\": %.4f, \"mean_ms\": %.4f, \"keys_per_s\": %.6g"
                ", \"offered_keys_per_s\": %.6g, \"batches\": %" PRId64 ", \"mean_batch_requests\": %.3f"
                ", \"max_inflight_batches\": %" PRId64 ", \"peak_pool_bytes\": %" PRId64 ", \"bypasses\": %" PRId64
                ", \"late_submissions\": %" PRId64 ", \"ring_stalls\": %" PRId64 ", \"verified\": %" PRId64
                ", \"failures\": %" PRId64 "}\n",
                requests, st.completed, st.rejected, st.p50_s * 1e3, st.p99_s * 1e3, st.max_s * 1e3, st.mean_s * 1e3,
                st.keys_per_s, double(offered) / seconds, st.batches, st.mean_batch_requests, st.max_inflight_batches,
                st.peak_pool_bytes, st.bypasses, late, starved, verified.load(), failures.load());
    return failures.load() ? 1 : 0;
}
