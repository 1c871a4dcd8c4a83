// Open-loop load generator for the serving layer (cs_server_*).
//
// Requests arrive as a Poisson process of rate --rate per second for
// --seconds; request sizes are floor(exp(U[ln min, ln(max+1)))) clipped to
// [min, max]; elements are records {int64 key, int32 payload = index within
// the request} (--dtype rec64) with keys uniform over 64 bits, or other key
// types.  Each request is submitted at its scheduled arrival time with that
// time as its arrival stamp, so generator lateness counts as latency.
// Buffers come from a ring; a slot is reused only after its previous
// request completed.  Every --verify-every-th result is checked (sorted keys,
// payload permutation, key/payload association).  Prints one JSON line.
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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
    int slots = 4, ring = 1024, verify_every = 97;
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
        else if (k == "--seed") seed = std::strtoull(v, nullptr, 10);
        else { std::fprintf(stderr, "unknown option %s\n", k.c_str()); return 2; }
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
    std::vector<uint64_t> ticket(static_cast<size_t>(requests), 0);
    int64_t verified = 0, failures = 0, late = 0;
    const double t0 = cs_server_now() + 0.05;
    for (int64_t r = 0; r < requests; ++r) {
        const size_t slot = size_t(r % ring);
        if (r >= ring) {                         // reuse: previous occupant must be done
            const int64_t prev = r - ring;
            rc = cs_server_wait(srv, ticket[size_t(prev)], -1.0);
            if (rc != CS_OK) { ++failures; }
            else if (verify_every > 0 && prev % verify_every == 0) {
                ++verified;
                if (!verify(&in[slot * size_t(max_n)], &out[slot * size_t(max_n)], size[size_t(prev)])) ++failures;
            }
            cs_server_release(srv, ticket[size_t(prev)]);
        }
        Rec* a = &in[slot * size_t(max_n)];
        const int64_t n = size[size_t(r)];
        for (int64_t j = 0; j < n; ++j) a[j] = Rec{int64_t(mix(seed ^ (uint64_t(r) << 20) ^ uint64_t(j))), int32_t(j), 0};
        const double at = t0 + arrive[size_t(r)];
        double now = cs_server_now();
        if (now > at) ++late;
        while (now < at) {
            if (at - now > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
            now = cs_server_now();
        }
        rc = cs_server_submit(srv, a, &out[slot * size_t(max_n)], n, at, &ticket[size_t(r)]);
        if (rc != CS_OK) { std::fprintf(stderr, "submit: %s\n", cs_strerror(rc)); ++failures; }
    }
    for (int64_t r = requests > ring ? requests - ring : 0; r < requests; ++r) {
        rc = cs_server_wait(srv, ticket[size_t(r)], -1.0);
        if (rc != CS_OK) ++failures;
        else if (verify_every > 0 && r % verify_every == 0) {
            ++verified;
            const size_t slot = size_t(r % ring);
            if (!verify(&in[slot * size_t(max_n)], &out[slot * size_t(max_n)], size[size_t(r)])) ++failures;
        }
    }
    cs_server_stats st;
    cs_server_get_stats(srv, &st);
    cs_server_destroy(srv);
    std::printf("{\"requests\": %" PRId64 ", \"completed\": %" PRId64 ", \"rejected\": %" PRId64
                ", \"p50_ms\": %.4f, \"p99_ms\": %.4f, \"max_ms\": %.4f, \"mean_ms\": %.4f, \"keys_per_s\": %.6g"
                ", \"offered_keys_per_s\": %.6g, \"batches\": %" PRId64 ", \"mean_batch_requests\": %.3f"
                ", \"max_inflight_batches\": %" PRId64 ", \"peak_pool_bytes\": %" PRId64
                ", \"late_submissions\": %" PRId64 ", \"verified\": %" PRId64 ", \"failures\": %" PRId64 "}\n",
                requests, st.completed, st.rejected, st.p50_s * 1e3, st.p99_s * 1e3, st.max_s * 1e3, st.mean_s * 1e3,
                st.keys_per_s, double(offered) / seconds, st.batches, st.mean_batch_requests, st.max_inflight_batches,
                st.peak_pool_bytes, late, verified, failures);
    return failures ? 1 : 0;
}
