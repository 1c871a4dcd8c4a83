// cusort_bench: the reference's benchmark driver (bench.cu:44-136, 240-271,
// 587-660; BENCHMARKS.md:116-137) on MI355X through the cusort-hip C ABI.
// Same options, size sweep, distributions, per-trial validation and CSV
// schema, so the reference's plotting scripts read its output unchanged.
// The CUB and Thrust columns are hipCUB BlockMergeSort and rocThrust
// thrust::sort (bench/baselines.hip); `cuda_runtime`, `thrust_version` and
// `cub_version` hold the HIP runtime, rocThrust and hipCUB versions.
//
// Methods and their (N, threads) domain follow bench.cu:523-552: bitonic_warp
// for N <= 32; for N >= 32 every TestIfSorted x FullSM x ASyncCpy
// combination of v1/v2 (threads = W), v3/v4 (every multiple of 32 up to
// min(W, --max-threads)), v5 (powers of two) and v6 (powers of two with
// W / threads <= 16); then the baselines.  --mode large times Sort
// (cs_sort_partition), SortMerge (cs_sort_merge) and HybridSort
// (cs_sort_hybrid) on single arrays of 2^k keys.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cusort_hip.h"
#include "baselines.hpp"

namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void cs_check(int rc, const std::string& what) {
    if (rc != CS_OK) throw std::runtime_error(what + ": " + cs_strerror(rc));
}

struct Options {
    std::string mode = "small", type = "both", prefix;
    int min_size = 1, max_size = 0, max_threads = 1024, arrays = 4096, batch_elements = 1 << 22, repetitions = 5;
    bool check_only = false;
    std::set<std::string> methods;
    bool enabled(const std::string& m) const { return methods.empty() || methods.count(m) != 0; }
};

std::string timestamp() {
    const auto now = std::time(nullptr);
    char v[32];
    std::strftime(v, sizeof(v), "%Y-%m-%d-%H-%M-%S", std::localtime(&now));
    return v;
}

void usage() {
    std::cout << "Usage: cusort_bench [--mode small|large|both] [--type int|record|both]\n"
                 "  --min-size N       Minimum array length (default 1)\n"
                 "  --max-size N       Maximum array length (default: bitonic capacity; 2^24 in large mode)\n"
                 "  --max-threads T    Sweep warp multiples up to T (default 1024)\n"
                 "  --arrays A         Maximum independent arrays per small batch (default 4096)\n"
                 "  --batch-elements M Bound N*A per batch (default 4194304)\n"
                 "  --repetitions R    Timed repetitions after warm-up (default 5)\n"
                 "  --output PREFIX    CSV path prefix\n"
                 "  --methods LIST     Comma-separated CSV method names (default: all)\n"
                 "  --check-only       Validate every applicable configuration with 3 arrays, no timing;\n"
                 "                     adds the EXTREMES distribution.\n"
                 "Small CSV: E_s is HIP-event batch time, X_s = E_s/(A*N*log2(N)); E_wall_s includes host\n"
                 "dispatch and completion; N=1 has X_s=nan.  cub_block_merge is hipCUB BlockMergeSort,\n"
                 "thrust_segment_loop calls rocThrust thrust::sort once per array.\n";
}

Options parse(int argc, char** argv) {
    Options o;
    o.prefix = "benchmark_" + timestamp();
    for (int arg = 1; arg < argc; ++arg) {
        const std::string key = argv[arg];
        if (key == "--help" || key == "-h") { usage(); std::exit(0); }
        if (key == "--check-only") { o.check_only = true; continue; }
        if (arg + 1 >= argc) throw std::runtime_error("Missing value for " + key);
        const std::string value = argv[++arg];
        if (key == "--mode") o.mode = value;
        else if (key == "--type") o.type = value;
        else if (key == "--output") o.prefix = value;
        else if (key == "--methods") {
            std::istringstream names(value);
            std::string m;
            while (std::getline(names, m, ',')) {
                if (m.empty()) throw std::runtime_error("Empty name in --methods");
                o.methods.insert(m);
            }
            if (o.methods.empty() || value.back() == ',') throw std::runtime_error("Empty name in --methods");
        } else {
            size_t used = 0;
            const long long v = std::stoll(value, &used);
            if (used != value.size() || v < 1 || v > (1 << 28)) throw std::runtime_error("Invalid positive integer for " + key);
            const int n = int(v);
            if (key == "--min-size") o.min_size = n;
            else if (key == "--max-size") o.max_size = n;
            else if (key == "--max-threads") o.max_threads = n;
            else if (key == "--arrays") o.arrays = n;
            else if (key == "--batch-elements") o.batch_elements = n;
            else if (key == "--repetitions") o.repetitions = n;
            else throw std::runtime_error("Unknown option: " + key);
        }
    }
    if (o.mode != "small" && o.mode != "large" && o.mode != "both") throw std::runtime_error("Invalid --mode");
    if (o.type != "int" && o.type != "record" && o.type != "both") throw std::runtime_error("Invalid --type");
    if (o.max_threads < 32 || o.max_threads > 1024 || o.max_threads % 32)
        throw std::runtime_error("--max-threads must be a multiple of 32 in [32,1024]");
    if (o.max_size && o.min_size > o.max_size) throw std::runtime_error("--min-size exceeds --max-size");
    const std::set<std::string> known = {"bitonic_warp", "bitonic_v1", "bitonic_v2", "bitonic_v3", "bitonic_v4_xor",
                                         "bitonic_v5_static", "bitonic_v6_register", "cub_block_merge",
                                         "thrust_segment_loop", "large_partition", "large_merge", "large_hybrid"};
    for (const auto& m : o.methods)
        if (!known.count(m)) throw std::runtime_error("Unknown --methods entry: " + m);
    if (o.check_only) { o.arrays = std::min(o.arrays, 3); o.repetitions = 0; }
    return o;
}

// ------------------------------------------------------------ inputs
// The reference's benchmark distributions (valgen.hpp:7-14, bench.cu:25-40,
// 158-192), generated on the host: SORTED, REVERSED, RANDOM, NORMAL,
// CONSTANT, BINARY, PERTURBED_2 / _8 / _32, and EXTREMES for --check-only.
enum Dist { SORTED, REVERSED, RANDOM, NORMAL, CONSTANT, BINARY, PERTURBED_2, PERTURBED_8, PERTURBED_32, EXTREMES, NDIST };
const char* kDistNames[] = {"SORTED", "REVERSED", "RANDOM", "NORMAL", "CONSTANT", "BINARY", "PERTURBED_2",
                            "PERTURBED_8", "PERTURBED_32", "EXTREMES"};

uint32_t mix_bits(uint32_t x) {   // the reference's integer hash (bench.cu:152-155)
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU;
    return x ^ (x >> 16);
}

int amplitude_of(int d) { return d == PERTURBED_2 ? 2 : d == PERTURBED_8 ? 8 : d == PERTURBED_32 ? 32 : 0; }

std::vector<int32_t> make_keys(int n, int count, int d) {
    std::vector<int32_t> k(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) {
        const int local = i % n;
        const uint32_t r = mix_bits(uint32_t(i) + 1234U);
        int key = 0;
        switch (d) {
            case SORTED: case PERTURBED_2: case PERTURBED_8: case PERTURBED_32: key = local; break;
            case REVERSED: key = n - 1 - local; break;
            case RANDOM: key = int(r % 1000001U); break;
            case NORMAL: {   // Box-Muller on the hash, clamped to [0, n]
                const double u = (double(r & 0xffffffU) + 1.0) / 16777217.0;
                const double v = double(mix_bits(r) & 0xffffffU) / 16777216.0;
                const double z = std::sqrt(-2.0 * std::log(u)) * std::cos(6.28318530718 * v);
                key = int(z * (1.0 + n / 4.0) + n / 2.0);
                key = key < 0 ? 0 : (key > n ? n : key);
                break;
            }
            case CONSTANT: key = 99; break;
            case BINARY: key = local % 2; break;
            default: {
                const int ex[] = {INT_MAX, INT_MIN, 0, INT_MAX, -1, 1};
                key = ex[local % 6];
            }
        }
        k[size_t(i)] = key;
    }
    if (const int amp = amplitude_of(d)) {   // bench_distributions.hpp:11-19, seeded per array
        const int delta = 4 * amp;
        for (int a = 0; a < count / n; ++a) {
            std::mt19937 rng(mix_bits(unsigned(a) + 1234U));
            std::uniform_int_distribution<int> noise(-delta, delta);
            for (int i = 0; i < n; ++i) k[size_t(a) * n + i] = 4 * i + noise(rng);
        }
    }
    return k;
}

// ------------------------------------------------------------ CSV
struct Config {
    std::string method;
    int testsorted = -1, fullsm = -1, async = -1;
    int threads = 0, items = 0, working = 0;
    size_t shared_bytes = 0;
    int registers = -1, active = -1;
    long long local_bytes = -1;
};

struct Csv {
    std::ofstream out;
    int runtime = 0;
    explicit Csv(const std::string& path) {
        const auto parent = std::filesystem::path(path).parent_path();
        if (!parent.empty()) std::filesystem::create_directories(parent);
        out.open(path);
        if (!out) throw std::runtime_error("Cannot open " + path);
        hip_check(hipRuntimeGetVersion(&runtime), "hipRuntimeGetVersion");
        out << "method,distribution,N,arrays,total_values,working_values,threads,items_per_thread,"
               "testsorted,fullsm,asynccpy,shared_bytes,trial,E_s,X_s,E_wall_s,"
               "wall_X_s,time_per_array_s,time_per_value_s,validated,cuda_runtime,thrust_version,cub_version,"
               "registers_per_thread,active_blocks_per_sm,local_bytes_per_thread\n";
        out << std::setprecision(17);
    }
    void write(const Config& c, int dist, int n, int arrays, int trial, double e, double wall) {
        const double values = double(n) * arrays;
        const double denom = n > 1 ? values * std::log2(double(n)) : 0;
        const double nan = std::numeric_limits<double>::quiet_NaN();
        out << c.method << ',' << kDistNames[dist] << ',' << n << ',' << arrays << ',' << size_t(n) * arrays << ','
            << c.working << ',' << c.threads << ',' << c.items << ',' << c.testsorted << ',' << c.fullsm << ','
            << c.async << ',' << c.shared_bytes << ',' << trial << ',' << e << ',' << (denom ? e / denom : nan)
            << ',' << wall << ',' << (denom ? wall / denom : nan) << ',' << e / arrays << ',' << e / values
            << ",1," << runtime << ',' << csbench::thrust_version() << ',' << csbench::hipcub_version() << ','
            << c.registers << ',' << c.active << ',' << c.local_bytes << '\n';
        out.flush();
    }
};

// ------------------------------------------------------------ batches
struct Batch {
    int type, n, arrays, count, dist;
    size_t esz;
    const Options& o;
    Csv& csv;
    void* input = nullptr;
    void* values = nullptr;
    std::vector<int32_t> in_host;      // interleaved key/payload for records
    std::vector<int32_t> expect_keys;  // sorted keys per array
    hipEvent_t b = nullptr, e = nullptr;

    Batch(int ty, int len, int arr, int d, const Options& opt, Csv& c)
        : type(ty), n(len), arrays(arr), count(len * arr), dist(d), esz(ty == 0 ? 4 : 8), o(opt), csv(c) {
        const std::vector<int32_t> keys = make_keys(n, count, d);
        if (type == 0) {
            in_host = keys;
        } else {
            in_host.resize(size_t(count) * 2);
            for (int i = 0; i < count; ++i) { in_host[2 * size_t(i)] = keys[size_t(i)]; in_host[2 * size_t(i) + 1] = i; }
        }
        expect_keys = keys;
        for (int a = 0; a < arrays; ++a)
            std::sort(expect_keys.begin() + size_t(a) * n, expect_keys.begin() + size_t(a + 1) * n);
        hip_check(hipMalloc(&input, size_t(count) * esz), "hipMalloc");
        hip_check(hipMalloc(&values, size_t(count) * esz), "hipMalloc");
        hip_check(hipMemcpy(input, in_host.data(), size_t(count) * esz, hipMemcpyHostToDevice), "upload");
        hip_check(hipEventCreate(&b), "event");
        hip_check(hipEventCreate(&e), "event");
    }
    ~Batch() {
        (void)hipFree(input); (void)hipFree(values);
        (void)hipEventDestroy(b); (void)hipEventDestroy(e);
    }
    void restore() {
        hip_check(hipMemcpy(values, input, size_t(count) * esz, hipMemcpyDeviceToDevice), "restore");
        hip_check(hipDeviceSynchronize(), "sync");
    }
    // bench.cu:209-228: keys equal the CPU sort; records keep their origin
    // in range, in the same array, equal to the input record, each once.
    void verify(const Config& c) {
        std::vector<int32_t> got(size_t(count) * (type == 0 ? 1 : 2));
        hip_check(hipMemcpy(got.data(), values, size_t(count) * esz, hipMemcpyDeviceToHost), "download");
        std::vector<char> seen(type == 0 ? 0 : size_t(count), 0);
        for (int i = 0; i < count; ++i) {
            bool ok = (type == 0 ? got[size_t(i)] : got[2 * size_t(i)]) == expect_keys[size_t(i)];
            if (ok && type == 1) {
                const int origin = got[2 * size_t(i) + 1];
                if (origin < 0 || origin >= count || origin / n != i / n) ok = false;
                else if (in_host[2 * size_t(origin)] != got[2 * size_t(i)] || in_host[2 * size_t(origin) + 1] != origin) ok = false;
                else if (seen[size_t(origin)]++) ok = false;
            }
            if (!ok)
                throw std::runtime_error("Validation failed: " + c.method + " N=" + std::to_string(n) + " threads=" +
                                         std::to_string(c.threads) + " sorted=" + std::to_string(c.testsorted) +
                                         " fullsm=" + std::to_string(c.fullsm) + " async=" + std::to_string(c.async) +
                                         " distribution=" + kDistNames[dist]);
        }
    }
    void run(const Config& c, const std::function<void()>& launch) {
        if (!o.enabled(c.method)) return;
        restore();   // warm-up and the full check stay outside the timed samples
        launch();
        hip_check(hipGetLastError(), "launch");
        hip_check(hipDeviceSynchronize(), "sync");
        verify(c);
        if (o.check_only) {
            std::cout << "PASS " << c.method << " N=" << n << " threads=" << c.threads << " sorted=" << c.testsorted
                      << " fullsm=" << c.fullsm << " async=" << c.async << ' ' << kDistNames[dist] << '\n';
            return;
        }
        for (int trial = 0; trial < o.repetitions; ++trial) {
            restore();   // never time the sorted output of a preceding run
            const auto t0 = std::chrono::steady_clock::now();
            hip_check(hipEventRecord(b, nullptr), "record");
            launch();
            hip_check(hipGetLastError(), "launch");
            hip_check(hipEventRecord(e, nullptr), "record");
            hip_check(hipEventSynchronize(e), "sync");
            const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, b, e), "elapsed");
            verify(c);   // validate this measured output, outside both timers
            csv.write(c, dist, n, arrays, trial, double(ms) * 1e-3, wall);
        }
    }
};

int next_pow2(int n) { int w = 1; while (w < n) w <<= 1; return w; }

const char* kMethodNames[] = {"auto", "bitonic_warp", "bitonic_v1", "bitonic_v2", "bitonic_v3", "bitonic_v4_xor",
                              "bitonic_v5_static", "bitonic_v6_register"};

void run_method(Batch& batch, int method, int threads, int working, int ts, int fs, int as) {
    const int dt = batch.type == 0 ? CS_I32 : CS_REC_I32_I32;
    const uint32_t flags = (ts ? CS_FLAG_TESTSORTED : 0) | (fs ? CS_FLAG_FULLSM : 0) | (as == 1 ? CS_FLAG_ASYNC : 0);
    int32_t bt = 0, ipt = 0, lds = 0, vgprs = -1;
    if (cs_batched_describe(dt, batch.n, method, threads, flags, &bt, &ipt, &lds, &vgprs, nullptr) != CS_OK) return;
    Config c{kMethodNames[method], ts, fs, as, threads ? threads : bt, ipt, working, size_t(lds), vgprs};
    batch.run(c, [&] {
        cs_check(cs_batched_sort(dt, batch.values, batch.values, batch.arrays, batch.n, method, threads, flags, nullptr),
                 c.method);
    });
}

void run_small_bitonic(Batch& batch) {
    const int working = std::max(32, next_pow2(batch.n));
    const int max_threads = batch.o.max_threads;
    for (int ts = 0; ts < 2; ++ts)
        for (int fs = 0; fs < 2; ++fs) {
            if (batch.n <= 32) run_method(batch, CS_BITONIC_WARP, 32, 32, ts, fs, -1);
            if (batch.n < 32) continue;
            for (int as = 0; as < 2; ++as) {
                if (working <= max_threads) {
                    run_method(batch, CS_BITONIC_V1, working, working, ts, fs, as);
                    run_method(batch, CS_BITONIC_V2, working, working, ts, fs, as);
                }
                for (int t = 32; t <= std::min(working, max_threads); t += 32) {
                    run_method(batch, CS_BITONIC_V3, t, working, ts, fs, as);
                    run_method(batch, CS_BITONIC_V4_XOR, t, working, ts, fs, as);
                }
                for (int t = 32; t <= std::min(working, max_threads); t *= 2) {
                    run_method(batch, CS_BITONIC_V5_STATIC, t, working, ts, fs, as);
                    if (working / t <= 16) run_method(batch, CS_BITONIC_V6_REGISTER, t, working, ts, fs, as);
                }
            }
        }
}

void run_block_merge(Batch& batch) {
    const int working = std::max(32, next_pow2(batch.n));
    for (const auto& cfg : csbench::block_merge_configs(batch.type)) {
        if (cfg.threads * cfg.items != working || cfg.threads > batch.o.max_threads) continue;
        csbench::KernelAttrs attrs;
        if (!csbench::block_merge_sort(batch.type, cfg.threads, cfg.items, batch.values, batch.n, 0, nullptr, &attrs))
            continue;
        Config c{"cub_block_merge", -1, -1, -1, cfg.threads, cfg.items, working, attrs.shared_bytes,
                 attrs.registers, attrs.active_blocks_per_cu, attrs.local_bytes};
        batch.run(c, [&] {
            csbench::block_merge_sort(batch.type, cfg.threads, cfg.items, batch.values, batch.n, batch.arrays, nullptr,
                                      nullptr);
        });
    }
}

void run_thrust(Batch& batch) {
    Config c{"thrust_segment_loop", -1, -1, -1, 0, 0, batch.n, 0};
    batch.run(c, [&] { csbench::thrust_segment_loop(batch.type, batch.values, batch.n, batch.arrays, nullptr); });
}

std::vector<int> small_sizes(int lo, int hi) {   // bench.cu:587-598
    std::set<int> s;
    for (int n = 1; n <= std::min(32, hi); ++n) s.insert(n);
    for (int n = 64; n <= std::min(1024, hi); n += 32) s.insert(n);
    for (int n = 32; n <= hi; n *= 2)
        for (int d : {-1, 0, 1})
            if (n + d >= 1 && n + d <= hi) s.insert(n + d);
    s.insert(hi);
    std::vector<int> r;
    for (int n : s) if (n >= lo) r.push_back(n);
    return r;
}

int safe_arrays(int n, size_t esz, const Options& o) {
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t cap = std::min(size_t(o.batch_elements), free_b / (8 * esz));
    if (cap < size_t(n)) throw std::runtime_error("Batch/memory budget cannot hold one array; increase --batch-elements");
    return int(std::min(size_t(o.arrays), cap / size_t(n)));
}

void bench_small(const Options& o, int type, const std::string& tname) {
    const int limit = cs_batched_max_n(type == 0 ? CS_I32 : CS_REC_I32_I32);
    const int maximum = o.max_size ? std::min(o.max_size, limit) : limit;
    if (o.mode == "small" && o.max_size > limit)
        throw std::runtime_error("--max-size exceeds the small bitonic capacity for " + tname + " (" + std::to_string(limit) + ")");
    if (o.min_size > maximum) return;
    Csv csv(o.prefix + "_" + tname + "_small.csv");
    for (int n : small_sizes(o.min_size, maximum)) {
        const int arrays = safe_arrays(n, type == 0 ? 4 : 8, o);
        for (int d = 0; d < PERTURBED_32 + 1 + (o.check_only ? 1 : 0); ++d) {
            std::cout << "small " << tname << " N=" << n << " arrays=" << arrays << ' ' << kDistNames[d] << std::endl;
            Batch batch(type, n, arrays, d, o, csv);
            run_small_bitonic(batch);
            run_block_merge(batch);
            run_thrust(batch);
        }
    }
}

void bench_large(const Options& o, int type, const std::string& tname) {
    const int dt = type == 0 ? CS_I32 : CS_REC_I32_I32;
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t esz = type == 0 ? 4 : 8;
    const int mem_limit = int(std::min(size_t(1) << 28, free_b / (8 * esz)));
    const int maximum = std::min(o.max_size ? o.max_size : (1 << 24), mem_limit);
    const int minimum = std::max(o.min_size, cs_batched_max_n(dt) + 1);
    Csv csv(o.prefix + "_" + tname + "_large.csv");
    for (int n = next_pow2(minimum); n <= maximum; n *= 2) {
        for (int d = 0; d < PERTURBED_32 + 1 + (o.check_only ? 1 : 0); ++d) {
            std::cout << "large " << tname << " N=" << n << ' ' << kDistNames[d] << std::endl;
            Batch batch(type, n, 1, d, o, csv);
            for (int ts = 0; ts < 2; ++ts)
                for (int fs = 0; fs < 2; ++fs) {
                    for (int as = 0; as < 2; ++as) {
                        const uint32_t flags = (ts ? CS_FLAG_TESTSORTED : 0) | (fs ? CS_FLAG_FULLSM : 0) | (as ? CS_FLAG_ASYNC : 0);
                        Config c{"large_partition", ts, fs, as, 0, 0, n, 0};
                        batch.run(c, [&] { cs_check(cs_sort_partition(dt, batch.values, n, 0, 0, flags, nullptr), c.method); });
                    }
                    Config m{"large_merge", ts, fs, -1, 0, 0, n, 0};
                    batch.run(m, [&] { cs_check(cs_sort_merge(dt, batch.values, n, 0, 0, nullptr, 0, 1, nullptr), m.method); });
                    Config h{"large_hybrid", ts, fs, -1, 0, 0, n, 0};
                    batch.run(h, [&] { cs_check(cs_sort_hybrid(dt, batch.values, n, nullptr, nullptr), h.method); });
                }
            run_thrust(batch);
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const Options o = parse(argc, argv);
        hip_check(hipSetDevice(0), "hipSetDevice");
        for (int type = 0; type < 2; ++type) {
            const std::string tname = type == 0 ? "int" : "record";
            if (o.type != "both" && o.type != tname) continue;
            if (o.mode != "large") bench_small(o, type, tname);
            if (o.mode != "small") bench_large(o, type, tname);
        }
    } catch (const std::exception& ex) {
        std::cerr << "Error: " << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
