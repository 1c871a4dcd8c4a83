// CPU oracle for the cusort-hip sorting path -- TEST INFRASTRUCTURE ONLY.
// See oracle.h for the contract, the reference citations and how each
// function is pinned.  Nothing in the product links or loads this file.
#include "oracle.h"

#include <parallel/algorithm>
#include <omp.h>
#include <sched.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

struct RecI32 { int32_t key; int32_t payload; };                 // IntAndData layout
struct RecI64 { int64_t key; int32_t payload; int32_t pad; };    // 16 B record

// ---------------------------------------------------------------- hashes --
inline uint64_t splitmix64(uint64_t x) {
    // First output of a SplitMix64 generator whose state is x (Steele, Lea,
    // Flood 2014; Vigna's splitmix64.c).
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The reference's stateless input hash (bench.cu:152-155), Wellons' lowbias32.
inline uint32_t mixbits(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU;
    return x ^ (x >> 16);
}

inline uint32_t u32hash(uint64_t seed, uint64_t g) { return uint32_t(splitmix64(seed ^ g) >> 32); }

constexpr uint64_t kSaltDup = 0x5bd1e9955bd1e995ull;
constexpr uint64_t kSaltSpecial = 0x00abcdef12345678ull;
constexpr uint64_t kSaltRun = 0x52554e5352554e53ull;
constexpr uint64_t kSaltRunBase = 0x4241534542415345ull;
constexpr uint64_t kSaltNormal = 0x4e4f524d4e4f524dull;

const int32_t kDistinct16[16] = {INT32_MIN, -1000000000, -65536, -1000, -7, -1, 0, 1,
                                 2, 3, 42, 1000, 65536, 1000000000, INT32_MAX - 1, INT32_MAX};
const uint32_t kSpecialF32[8] = {0x7FC00000u, 0xFFC00000u, 0x7F800001u, 0xFFFFFFFFu,
                                 0x80000000u, 0x7F800000u, 0xFF800000u, 0x00000000u};

// int32 value of distribution `dist` at global index g (local position p of
// an array of length n, array index a).
int32_t dist_i32(int dist, int64_t n, int64_t g, uint64_t seed) {
    const int64_t p = g % n;
    const int64_t a = g / n;
    switch (dist) {
        case ORC_D_UNIFORM: return int32_t(u32hash(seed, uint64_t(g)));
        case ORC_D_SORTED: return int32_t(p);
        case ORC_D_REVERSED: return int32_t(n - 1 - p);
        case ORC_D_CONSTANT: return 99;                                  // bench.cu:182
        case ORC_D_DISTINCT16: return kDistinct16[u32hash(seed, uint64_t(g)) & 15u];
        case ORC_D_PERTURBED:
            // 4*p + U{-8..8}: the spacing 4 and amplitude 2 spacings of the
            // reference's PERTURBED_2 (bench_distributions.hpp:11-19), drawn from a
            // stateless hash instead of mt19937 so host and device agree.
            return int32_t(4 * p + int64_t(u32hash(seed, uint64_t(g)) % 17u) - 8);
        case ORC_D_ORGAN_PIPE: return int32_t(std::min<int64_t>(p, n - 1 - p));
        case ORC_D_ALT_EXTREMES: return (p & 1) ? INT32_MAX : INT32_MIN;
        case ORC_D_DUP50: {
            const int64_t half = (n + 1) / 2;
            const int64_t q = p < half ? p : int64_t(u32hash(seed ^ kSaltDup, uint64_t(g)) % uint64_t(half));
            return int32_t(u32hash(seed, uint64_t(a * n + q)));
        }
        case ORC_D_EXTREMES6: {                                           // bench.cu:184-187
            static const int32_t e[6] = {INT32_MAX, INT32_MIN, 0, INT32_MAX, -1, 1};
            return e[p % 6];
        }
        case ORC_D_RANDOM1M: return int32_t(mixbits(uint32_t(g) + 1234u) % 1000001u); // bench.cu:171
        case ORC_D_BINARY: return int32_t(p % 2);                          // bench.cu:183
        case ORC_D_PRESORTED_RUNS: {
            const uint64_t block = uint64_t(g) / 4096u;
            if (u32hash(seed ^ kSaltRun, block) % 10u == 0) {
                const int32_t base = int32_t(u32hash(seed ^ kSaltRunBase, block) >> 1) - (1 << 30);
                return base + int32_t(uint64_t(g) % 4096u);
            }
            return int32_t(u32hash(seed, uint64_t(g)));
        }
        default: return 0;
    }
}

uint64_t dist_u64(int dist, int64_t n, int64_t g, uint64_t seed) {
    if (dist == ORC_D_UNIFORM) return splitmix64(seed ^ uint64_t(g));
    if (dist == ORC_D_PRESORTED_RUNS) {
        const uint64_t block = uint64_t(g) / 4096u;
        if (u32hash(seed ^ kSaltRun, block) % 10u == 0)
            return (splitmix64(seed ^ kSaltRunBase ^ block) >> 1) + uint64_t(g) % 4096u;
        return splitmix64(seed ^ uint64_t(g));
    }
    return uint64_t(int64_t(dist_i32(dist, n, g, seed)));
}

uint32_t dist_f32_bits(int dist, int64_t n, int64_t g, uint64_t seed) {
    if (dist == ORC_D_UNIFORM) return u32hash(seed, uint64_t(g));   // every bit pattern
    if (dist == ORC_D_NORMAL_SPECIAL) {
        const uint32_t h = u32hash(seed ^ kSaltSpecial, uint64_t(g));
        if (h % 100u == 0) return kSpecialF32[(h >> 8) & 7u];
        // Irwin-Hall(12) - 6 approximates N(0,1); integer sum, one rounding.
        int64_t sum = 0;
        for (uint64_t i = 0; i < 12; ++i) sum += int64_t(u32hash(seed ^ kSaltNormal, uint64_t(g) * 16u + i));
        sum -= int64_t(6) << 32;
        const float z = float(double(sum) * 0x1p-32);
        uint32_t bits; std::memcpy(&bits, &z, 4); return bits;
    }
    const float v = float(dist_i32(dist, n, g, seed));
    uint32_t bits; std::memcpy(&bits, &v, 4); return bits;
}

const uint64_t kSpecialF64[8] = {0x7FF8000000000000ull, 0xFFF8000000000000ull, 0x7FF0000000000001ull,
                                 0xFFFFFFFFFFFFFFFFull, 0x8000000000000000ull, 0x7FF0000000000000ull,
                                 0xFFF0000000000000ull, 0ull};   // qNaN, -qNaN, sNaN, -NaN, -0, +inf, -inf, +0

uint64_t dist_f64_bits(int dist, int64_t n, int64_t g, uint64_t seed) {
    if (dist == ORC_D_UNIFORM) return splitmix64(seed ^ uint64_t(g));   // every bit pattern
    if (dist == ORC_D_NORMAL_SPECIAL) {
        const uint32_t h = u32hash(seed ^ kSaltSpecial, uint64_t(g));
        if (h % 100u == 0) return kSpecialF64[(h >> 8) & 7u];
        int64_t sum = 0;   // Irwin-Hall(12) - 6, exact in double
        for (uint64_t i = 0; i < 12; ++i) sum += int64_t(u32hash(seed ^ kSaltNormal, uint64_t(g) * 16u + i));
        sum -= int64_t(6) << 32;
        const double z = double(sum) * 0x1p-32;
        uint64_t bits; std::memcpy(&bits, &z, 8); return bits;
    }
    const double v = double(dist_i32(dist, n, g, seed));
    uint64_t bits; std::memcpy(&bits, &v, 8); return bits;
}

// ------------------------------------------------------------ key orders --
inline uint32_t f32_ordered(uint32_t b) { return b ^ ((b >> 31) ? 0xFFFFFFFFu : 0x80000000u); }
inline uint32_t f32_unordered(uint32_t k) { return k ^ ((k >> 31) ? 0x80000000u : 0xFFFFFFFFu); }
inline uint64_t f64_ordered(uint64_t b) { return b ^ ((b >> 63) ? ~uint64_t(0) : (uint64_t(1) << 63)); }

template <int D> struct Traits;
template <> struct Traits<ORC_I32> {
    using T = int32_t; using K = int32_t;
    static K key(const T& v) { return v; }
    static bool same(const T& a, const T& b) { return a == b; }
};
template <> struct Traits<ORC_U32> {
    using T = uint32_t; using K = uint32_t;
    static K key(const T& v) { return v; }
    static bool same(const T& a, const T& b) { return a == b; }
};
template <> struct Traits<ORC_F32> {
    using T = uint32_t; using K = uint32_t;   // raw IEEE bits
    static K key(const T& v) { return f32_ordered(v); }
    static bool same(const T& a, const T& b) { return a == b; }
};
template <> struct Traits<ORC_F64> {
    using T = uint64_t; using K = uint64_t;   // raw IEEE bits
    static K key(const T& v) { return f64_ordered(v); }
    static bool same(const T& a, const T& b) { return a == b; }
};
template <> struct Traits<ORC_U64> {
    using T = uint64_t; using K = uint64_t;
    static K key(const T& v) { return v; }
    static bool same(const T& a, const T& b) { return a == b; }
};
template <> struct Traits<ORC_I64> {
    using T = int64_t; using K = int64_t;
    static K key(const T& v) { return v; }
    static bool same(const T& a, const T& b) { return a == b; }
};
template <> struct Traits<ORC_REC_I32_I32> {
    using T = RecI32; using K = int32_t;
    static K key(const T& v) { return v.key; }
    static bool same(const T& a, const T& b) { return a.key == b.key && a.payload == b.payload; }
};
template <> struct Traits<ORC_REC_I64_I32> {
    using T = RecI64; using K = int64_t;
    static K key(const T& v) { return v.key; }
    static bool same(const T& a, const T& b) { return a.key == b.key && a.payload == b.payload; }
};

template <int D> constexpr bool kHasPayload = (D == ORC_REC_I32_I32 || D == ORC_REC_I64_I32);

template <int D> void sort_range(typename Traits<D>::T* first, typename Traits<D>::T* last) {
    using Tr = Traits<D>;
    std::sort(first, last, [](const typename Tr::T& a, const typename Tr::T& b) { return Tr::key(a) < Tr::key(b); });
}

template <typename F> void dispatch(int dtype, F&& f) {
    switch (dtype) {
        case ORC_I32: f(std::integral_constant<int, ORC_I32>{}); break;
        case ORC_REC_I32_I32: f(std::integral_constant<int, ORC_REC_I32_I32>{}); break;
        case ORC_F32: f(std::integral_constant<int, ORC_F32>{}); break;
        case ORC_U64: f(std::integral_constant<int, ORC_U64>{}); break;
        case ORC_REC_I64_I32: f(std::integral_constant<int, ORC_REC_I64_I32>{}); break;
        case ORC_U32: f(std::integral_constant<int, ORC_U32>{}); break;
        case ORC_I64: f(std::integral_constant<int, ORC_I64>{}); break;
        case ORC_F64: f(std::integral_constant<int, ORC_F64>{}); break;
        default: break;
    }
}

bool valid_dtype(int d) { return d >= ORC_I32 && d <= ORC_F64; }

int host_cores() {
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) {
        const int c = CPU_COUNT(&set);
        if (c > 0) return c;
    }
    const unsigned h = std::thread::hardware_concurrency();
    return h ? int(h) : 1;
}

int resolve_threads(int threads) { return threads > 0 ? threads : host_cores(); }

// Split segments [0, segs) into `parts` contiguous ranges of roughly equal
// sum of n*log2(n) (the std::sort cost model).
std::vector<int64_t> balance(const int64_t* offsets, int64_t segs, int parts) {
    std::vector<double> prefix(size_t(segs) + 1, 0.0);
    for (int64_t s = 0; s < segs; ++s) {
        const double len = double(offsets[s + 1] - offsets[s]);
        prefix[size_t(s) + 1] = prefix[size_t(s)] + (len > 1 ? len * std::log2(len) : len);
    }
    std::vector<int64_t> cut(size_t(parts) + 1, segs);
    cut[0] = 0;
    for (int p = 1; p < parts; ++p) {
        const double target = prefix[size_t(segs)] * p / parts;
        cut[size_t(p)] = std::lower_bound(prefix.begin(), prefix.end(), target) - prefix.begin();
        cut[size_t(p)] = std::min<int64_t>(std::max<int64_t>(cut[size_t(p)], cut[size_t(p) - 1]), segs);
    }
    cut[size_t(parts)] = segs;
    return cut;
}

template <int D>
void sort_segments(typename Traits<D>::T* data, const int64_t* offsets, int64_t segs, int threads) {
    threads = std::max(1, std::min<int>(threads, int(std::max<int64_t>(segs, 1))));
    const auto cut = balance(offsets, segs, threads);
    auto work = [&](int part) {
        for (int64_t s = cut[size_t(part)]; s < cut[size_t(part) + 1]; ++s)
            sort_range<D>(data + offsets[s], data + offsets[s + 1]);
    };
    if (threads == 1) { work(0); return; }
    std::vector<std::thread> pool;
    for (int p = 0; p < threads; ++p) pool.emplace_back(work, p);
    for (auto& t : pool) t.join();
}

std::vector<int64_t> uniform_offsets(int64_t n, int64_t arrays) {
    std::vector<int64_t> off(size_t(arrays) + 1);
    for (int64_t a = 0; a <= arrays; ++a) off[size_t(a)] = a * n;
    return off;
}

void put_msg(char* msg, int len, const char* text) {
    if (msg && len > 0) { std::snprintf(msg, size_t(len), "%s", text); }
}

template <int D>
int64_t validate(const typename Traits<D>::T* input, const typename Traits<D>::T* result,
                 const int64_t* offsets, int64_t segs, char* msg, int msglen) {
    using Tr = Traits<D>;
    using T = typename Tr::T;
    const int64_t count = offsets[segs];
    std::vector<T> expected(input, input + count);
    sort_segments<D>(expected.data(), offsets, segs, host_cores());
    std::vector<uint8_t> seen(kHasPayload<D> ? size_t(count) : 0, 0);
    char buf[256];
    for (int64_t s = 0; s < segs; ++s) {
        for (int64_t i = offsets[s]; i < offsets[s + 1]; ++i) {
            // Bit-exact key sequence (bench.cu:204-207 compares keys by order;
            // here raw key bits so -0/+0 and NaN payloads must match too).
            if (!(Tr::key(result[i]) == Tr::key(expected[size_t(i)]))) {
                std::snprintf(buf, sizeof buf, "key mismatch at %lld (segment %lld)", (long long)i, (long long)s);
                put_msg(msg, msglen, buf);
                return i + 1;
            }
            if constexpr (kHasPayload<D>) {
                const int64_t origin = int64_t(result[i].payload);
                if (origin < offsets[s] || origin >= offsets[s + 1]) {          // bench.cu:218
                    std::snprintf(buf, sizeof buf, "payload %lld out of segment %lld at %lld",
                                  (long long)origin, (long long)s, (long long)i);
                    put_msg(msg, msglen, buf);
                    return i + 1;
                }
                if (!Tr::same(input[origin], result[i])) {                      // bench.cu:220
                    std::snprintf(buf, sizeof buf, "record at %lld does not match input[%lld]",
                                  (long long)i, (long long)origin);
                    put_msg(msg, msglen, buf);
                    return i + 1;
                }
                if (seen[size_t(origin)]++) {                                    // bench.cu:221
                    std::snprintf(buf, sizeof buf, "payload %lld appears twice (at %lld)",
                                  (long long)origin, (long long)i);
                    put_msg(msg, msglen, buf);
                    return i + 1;
                }
            }
        }
    }
    put_msg(msg, msglen, "ok");
    return 0;
}

double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v.empty() ? 0.0 : v[v.size() / 2];
}

}  // namespace

extern "C" {

uint64_t orc_splitmix64(uint64_t x) { return splitmix64(x); }
uint32_t orc_mixbits(uint32_t x) { return mixbits(x); }
uint32_t orc_f32_to_ordered(uint32_t bits) { return f32_ordered(bits); }
uint32_t orc_ordered_to_f32(uint32_t key) { return f32_unordered(key); }
int orc_host_cores(void) { return host_cores(); }

int orc_generate(int dtype, int dist, int64_t n, int64_t count, uint64_t seed, void* out) {
    if (!valid_dtype(dtype) || dist < 0 || dist >= ORC_D_COUNT || n < 1 || count < 0 || !out) return -1;
    #pragma omp parallel for schedule(static) num_threads(host_cores())
    for (int64_t g = 0; g < count; ++g) {
        switch (dtype) {
            case ORC_I32: static_cast<int32_t*>(out)[g] = dist_i32(dist, n, g, seed); break;
            case ORC_U32: static_cast<uint32_t*>(out)[g] = uint32_t(dist_i32(dist, n, g, seed)); break;
            case ORC_F32: static_cast<uint32_t*>(out)[g] = dist_f32_bits(dist, n, g, seed); break;
            case ORC_U64: static_cast<uint64_t*>(out)[g] = dist_u64(dist, n, g, seed); break;
            case ORC_I64: static_cast<int64_t*>(out)[g] = int64_t(dist_u64(dist, n, g, seed)); break;
            case ORC_F64: static_cast<uint64_t*>(out)[g] = dist_f64_bits(dist, n, g, seed); break;
            case ORC_REC_I32_I32:
                static_cast<RecI32*>(out)[g] = RecI32{dist_i32(dist, n, g, seed), int32_t(g)};   // bench.cu:189
                break;
            case ORC_REC_I64_I32:
                static_cast<RecI64*>(out)[g] = RecI64{int64_t(dist_u64(dist, n, g, seed)), int32_t(g), 0};
                break;
        }
    }
    return 0;
}

int orc_sort_segmented(int dtype, const void* in, void* out, const int64_t* offsets, int64_t segs, int threads) {
    if (!valid_dtype(dtype) || !offsets || segs < 0) return -1;
    int rc = 0;
    dispatch(dtype, [&](auto d) {
        constexpr int D = decltype(d)::value;
        using T = typename Traits<D>::T;
        const int64_t count = offsets[segs];
        if (in != out) std::memcpy(out, in, size_t(count) * sizeof(T));
        sort_segments<D>(static_cast<T*>(out), offsets, segs, resolve_threads(threads));
    });
    return rc;
}

int orc_sort_batched(int dtype, const void* in, void* out, int64_t n, int64_t arrays, int threads) {
    if (n < 1 || arrays < 0) return -1;
    const auto off = uniform_offsets(n, arrays);
    return orc_sort_segmented(dtype, in, out, off.data(), arrays, threads);
}

int orc_sort_large(int dtype, const void* in, void* out, int64_t count, int threads) {
    if (!valid_dtype(dtype) || count < 0) return -1;
    dispatch(dtype, [&](auto d) {
        constexpr int D = decltype(d)::value;
        using Tr = Traits<D>;
        using T = typename Tr::T;
        T* data = static_cast<T*>(out);
        if (in != out) std::memcpy(out, in, size_t(count) * sizeof(T));
        omp_set_num_threads(resolve_threads(threads));
        __gnu_parallel::sort(data, data + count,
                             [](const T& a, const T& b) { return Tr::key(a) < Tr::key(b); });
    });
    return 0;
}

int64_t orc_validate_segmented(int dtype, const void* input, const void* result, const int64_t* offsets,
                               int64_t segs, char* msg, int msglen) {
    if (!valid_dtype(dtype) || !offsets || segs < 0 || !input || !result) return -1;
    int64_t rc = 0;
    dispatch(dtype, [&](auto d) {
        constexpr int D = decltype(d)::value;
        using T = typename Traits<D>::T;
        rc = validate<D>(static_cast<const T*>(input), static_cast<const T*>(result), offsets, segs, msg, msglen);
    });
    return rc;
}

int64_t orc_validate_batched(int dtype, const void* input, const void* result, int64_t n, int64_t arrays,
                             char* msg, int msglen) {
    if (n < 1 || arrays < 0) return -1;
    const auto off = uniform_offsets(n, arrays);
    return orc_validate_segmented(dtype, input, result, off.data(), arrays, msg, msglen);
}

double orc_time_segmented(int dtype, const void* in, const int64_t* offsets, int64_t segs, int threads,
                          int trials, int* threads_used) {
    if (!valid_dtype(dtype) || !offsets || segs < 0 || trials < 1) return -1.0;
    const int t = resolve_threads(threads);
    if (threads_used) *threads_used = t;
    std::vector<double> times;
    dispatch(dtype, [&](auto d) {
        constexpr int D = decltype(d)::value;
        using T = typename Traits<D>::T;
        const int64_t count = offsets[segs];
        std::vector<T> work(static_cast<size_t>(count));
        for (int r = 0; r < trials; ++r) {
            std::memcpy(work.data(), in, size_t(count) * sizeof(T));   // restore (bench.cu:335)
            const auto t0 = std::chrono::steady_clock::now();
            sort_segments<D>(work.data(), offsets, segs, t);
            times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
    });
    return median(times);
}

double orc_time_batched(int dtype, const void* in, int64_t n, int64_t arrays, int threads, int trials,
                        int* threads_used) {
    if (n < 1 || arrays < 0) return -1.0;
    const auto off = uniform_offsets(n, arrays);
    return orc_time_segmented(dtype, in, off.data(), arrays, threads, trials, threads_used);
}

double orc_time_large(int dtype, const void* in, int64_t count, int threads, int trials, int* threads_used) {
    if (!valid_dtype(dtype) || count < 0 || trials < 1) return -1.0;
    const int t = resolve_threads(threads);
    if (threads_used) *threads_used = t;
    std::vector<double> times;
    dispatch(dtype, [&](auto d) {
        constexpr int D = decltype(d)::value;
        using Tr = Traits<D>;
        using T = typename Tr::T;
        std::vector<T> work(static_cast<size_t>(count));
        omp_set_num_threads(t);
        for (int r = 0; r < trials; ++r) {
            std::memcpy(work.data(), in, size_t(count) * sizeof(T));
            const auto t0 = std::chrono::steady_clock::now();
            if (t == 1) std::sort(work.begin(), work.end(), [](const T& a, const T& b) { return Tr::key(a) < Tr::key(b); });
            else __gnu_parallel::sort(work.begin(), work.end(), [](const T& a, const T& b) { return Tr::key(a) < Tr::key(b); });
            times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
    });
    return median(times);
}

}  // extern "C"
