// cusort_bench: benchmark driver of the cusort-hip C ABI that accepts the
// reference bench's command line and writes its CSV schema, so that scripts
// written for the reference read MI355X results unchanged.
//
// Interface kept from the reference (bench.cu:44-136 options and messages,
// 240-271 columns, 587-613 size sweep and batch sizing, 152-192 inputs,
// 209-228 checks; BENCHMARKS.md:116-137).  Everything behind it is this
// project's: options come from a table, the configurations of a batch are
// the candidate (method, threads, flags) triples that the library itself
// accepts (cs_batched_describe), rows are written through a column table,
// and a Workload object owns the device buffers, the restore copy and the
// validation.  The CUB and Thrust columns are hipCUB BlockMergeSort and a
// rocThrust thrust::sort per array (bench/baselines.hip).
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
#include <iostream>
#include <limits>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cusort_hip.h"
#include "baselines.hpp"

namespace bench {

[[noreturn]] void fail(const std::string& what) { throw std::runtime_error(what); }
void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

// ================================================================ settings
struct Settings {
    std::string mode = "small", type = "both", prefix;
    long long min_size = 1, max_size = 0, max_threads = 1024, arrays = 4096, batch_elements = 1 << 22,
              repetitions = 5;
    bool check_only = false;
    std::set<std::string> methods;   // empty: every method
    bool wants(const std::string& m) const { return methods.empty() || methods.count(m) > 0; }
};

const std::set<std::string>& known_methods() {
    static const std::set<std::string> names = {
        "bitonic_warp", "bitonic_v1", "bitonic_v2", "bitonic_v3", "bitonic_v4_xor", "bitonic_v5_static",
        "bitonic_v6_register", "cub_block_merge", "thrust_segment_loop", "large_partition", "large_merge",
        "large_hybrid"};
    return names;
}

// One row per option: how it is spelled, what it takes, and what it sets.
enum class Arg { None, Text, Count };
struct OptionRow {
    const char* name;
    Arg arg;
    const char* help;
    std::function<void(Settings&, const std::string&, long long)> set;
};

std::vector<std::string> split_list(const std::string& text) {
    std::vector<std::string> parts;
    std::string cur;
    for (const char ch : text + ",") {
        if (ch != ',') { cur += ch; continue; }
        if (cur.empty()) fail("Empty name in --methods");
        parts.push_back(cur);
        cur.clear();
    }
    return parts;
}

const std::vector<OptionRow>& option_table() {
    static const std::vector<OptionRow> rows = {
        {"--mode", Arg::Text, "small|large|both (default small)",
         [](Settings& s, const std::string& v, long long) { s.mode = v; }},
        {"--type", Arg::Text, "int|record|both (default both)",
         [](Settings& s, const std::string& v, long long) { s.type = v; }},
        {"--min-size", Arg::Count, "Minimum array length (default 1)",
         [](Settings& s, const std::string&, long long n) { s.min_size = n; }},
        {"--max-size", Arg::Count, "Maximum array length (default: bitonic capacity; 2^24 in large mode)",
         [](Settings& s, const std::string&, long long n) { s.max_size = n; }},
        {"--max-threads", Arg::Count, "Largest thread count swept, a multiple of 32 (default 1024)",
         [](Settings& s, const std::string&, long long n) { s.max_threads = n; }},
        {"--arrays", Arg::Count, "Maximum independent arrays per small batch (default 4096)",
         [](Settings& s, const std::string&, long long n) { s.arrays = n; }},
        {"--batch-elements", Arg::Count, "Bound N*A per batch (default 4194304)",
         [](Settings& s, const std::string&, long long n) { s.batch_elements = n; }},
        {"--repetitions", Arg::Count, "Timed repetitions after warm-up (default 5)",
         [](Settings& s, const std::string&, long long n) { s.repetitions = n; }},
        {"--output", Arg::Text, "CSV path prefix",
         [](Settings& s, const std::string& v, long long) { s.prefix = v; }},
        {"--methods", Arg::Text, "Comma-separated CSV method names (default: all)",
         [](Settings& s, const std::string& v, long long) {
             for (const auto& m : split_list(v)) s.methods.insert(m);
         }},
        {"--check-only", Arg::None,
         "Validate every applicable configuration with 3 arrays, no timing; adds the EXTREMES distribution",
         [](Settings& s, const std::string&, long long) { s.check_only = true; }},
    };
    return rows;
}

void print_help() {
    std::cout << "Usage: cusort_bench [options]\n";
    for (const auto& r : option_table()) {
        std::string head = std::string("  ") + r.name + (r.arg == Arg::None ? "" : r.arg == Arg::Count ? " N" : " S");
        head.resize(std::max<size_t>(head.size() + 1, 22), ' ');
        std::cout << head << r.help << '\n';
    }
    std::cout << "Small CSV: E_s is the HIP-event time of one batch, X_s = E_s/(A*N*log2(N)) (nan for N=1);\n"
                 "E_wall_s adds host dispatch and completion.  cub_block_merge is hipCUB BlockMergeSort,\n"
                 "thrust_segment_loop calls rocThrust thrust::sort once per array.\n";
}

long long parse_count(const std::string& key, const std::string& text) {
    long long v = 0;
    size_t used = 0;
    try { v = std::stoll(text, &used); } catch (const std::exception&) { used = 0; }
    if (used == 0 || used != text.size() || v < 1 || v > (1LL << 28)) fail("Invalid positive integer for " + key);
    return v;
}

std::string now_stamp() {
    char buf[32];
    const std::time_t t = std::time(nullptr);
    std::strftime(buf, sizeof buf, "%Y-%m-%d-%H-%M-%S", std::localtime(&t));
    return buf;
}

Settings parse_command_line(int argc, char** argv) {
    Settings s;
    s.prefix = "benchmark_" + now_stamp();
    const auto& table = option_table();
    for (int i = 1; i < argc; ++i) {
        const std::string key = argv[i];
        if (key == "--help" || key == "-h") {
            print_help();
            std::exit(0);
        }
        const auto row = std::find_if(table.begin(), table.end(), [&](const OptionRow& r) { return key == r.name; });
        if (row == table.end()) {
            // unknown names still consume a value, as in the reference; they
            // are reported once that value is known to exist
            if (i + 1 >= argc) fail("Missing value for " + key);
            fail("Unknown option: " + key);
        }
        if (row->arg == Arg::None) {
            row->set(s, "", 0);
            continue;
        }
        if (i + 1 >= argc) fail("Missing value for " + key);
        const std::string value = argv[++i];
        row->set(s, value, row->arg == Arg::Count ? parse_count(key, value) : 0);
    }
    // constraints between options
    if (s.mode != "small" && s.mode != "large" && s.mode != "both") fail("Invalid --mode");
    if (s.type != "int" && s.type != "record" && s.type != "both") fail("Invalid --type");
    if (s.max_threads < 32 || s.max_threads > 1024 || s.max_threads % 32 != 0)
        fail("--max-threads must be a multiple of 32 in [32,1024]");
    if (s.max_size && s.min_size > s.max_size) fail("--min-size exceeds --max-size");
    for (const auto& m : s.methods)
        if (!known_methods().count(m)) fail("Unknown --methods entry: " + m);
    if (s.check_only) {
        s.arrays = std::min(s.arrays, 3LL);
        s.repetitions = 0;
    }
    return s;
}

// ================================================================ inputs
// The reference's benchmark distributions (bench.cu:158-192,
// bench_distributions.hpp:11-19); `hash` is the reference's integer hash of
// the global index + 1234 (bench.cu:152-155).
uint32_t hash32(uint32_t x) {
    x = (x ^ (x >> 16)) * 0x7feb352dU;
    x = (x ^ (x >> 15)) * 0x846ca68bU;
    return x ^ (x >> 16);
}

struct Distribution {
    const char* name;
    int perturb;        // PERTURBED_a: sorted positions 4 i + U{-4a .. 4a} per array
    bool check_only;    // only in --check-only sweeps
    std::function<int32_t(int local, int n, uint32_t hash)> key;
};

const std::vector<Distribution>& distributions() {
    static const std::vector<Distribution> table = {
        {"SORTED", 0, false, [](int local, int, uint32_t) { return local; }},
        {"REVERSED", 0, false, [](int local, int n, uint32_t) { return n - 1 - local; }},
        {"RANDOM", 0, false, [](int, int, uint32_t h) { return int32_t(h % 1000001U); }},
        {"NORMAL", 0, false,
         [](int, int n, uint32_t h) {   // Box-Muller on two hashes, clamped to [0, n]
             const double u = (double(h & 0xffffffU) + 1.0) / 16777217.0;
             const double v = double(hash32(h) & 0xffffffU) / 16777216.0;
             const double z = std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * v);
             return int32_t(std::clamp(int(z * (1.0 + n / 4.0) + n / 2.0), 0, n));
         }},
        {"CONSTANT", 0, false, [](int, int, uint32_t) { return 99; }},
        {"BINARY", 0, false, [](int local, int, uint32_t) { return local % 2; }},
        {"PERTURBED_2", 2, false, nullptr},
        {"PERTURBED_8", 8, false, nullptr},
        {"PERTURBED_32", 32, false, nullptr},
        {"EXTREMES", 0, true,
         [](int local, int, uint32_t) {
             static const int32_t cycle[6] = {INT_MAX, INT_MIN, 0, INT_MAX, -1, 1};
             return cycle[local % 6];
         }},
    };
    return table;
}

std::vector<int32_t> generate(const Distribution& d, int n, int arrays) {
    std::vector<int32_t> keys(size_t(n) * size_t(arrays));
    if (d.perturb) {
        const int reach = 4 * d.perturb;
        for (int a = 0; a < arrays; ++a) {
            std::mt19937 rng(hash32(uint32_t(a) + 1234U));
            std::uniform_int_distribution<int> jitter(-reach, reach);
            for (int i = 0; i < n; ++i) keys[size_t(a) * n + i] = 4 * i + jitter(rng);
        }
        return keys;
    }
    for (size_t i = 0; i < keys.size(); ++i) keys[i] = d.key(int(i % size_t(n)), n, hash32(uint32_t(i) + 1234U));
    return keys;
}

// ================================================================ workload
// One batch on the device: the pristine input, the working buffer that is
// restored before every launch, and what a correct output must be.
class Workload {
public:
    Workload(bool records, int n, int arrays, const Distribution& d)
        : records_(records), n_(n), arrays_(arrays), count_(size_t(n) * arrays), dist_(d.name) {
        keys_ = generate(d, n, arrays);
        host_.resize(count_ * words());
        for (size_t i = 0; i < count_; ++i) {
            host_[i * words()] = keys_[i];
            if (records_) host_[i * words() + 1] = int32_t(i);   // payload: global index (bench.cu:189)
        }
        sorted_ = keys_;
        for (int a = 0; a < arrays_; ++a)
            std::sort(sorted_.begin() + size_t(a) * n_, sorted_.begin() + size_t(a + 1) * n_);
        hip_ok(hipMalloc(&pristine_, bytes()), "hipMalloc");
        hip_ok(hipMalloc(&work_, bytes()), "hipMalloc");
        hip_ok(hipMemcpy(pristine_, host_.data(), bytes(), hipMemcpyHostToDevice), "upload");
    }
    ~Workload() {
        (void)hipFree(pristine_);
        (void)hipFree(work_);
    }
    Workload(const Workload&) = delete;
    Workload& operator=(const Workload&) = delete;

    void* data() const { return work_; }
    int n() const { return n_; }
    int arrays() const { return arrays_; }
    bool records() const { return records_; }
    const char* dist() const { return dist_; }
    int dtype() const { return records_ ? CS_REC_I32_I32 : CS_I32; }

    void restore() const {
        hip_ok(hipMemcpy(work_, pristine_, bytes(), hipMemcpyDeviceToDevice), "restore");
        hip_ok(hipDeviceSynchronize(), "restore");
    }
    // Empty when the working buffer holds a correct result: every key equals
    // the host sort of its array; every record names an input record of its
    // own array, carries that record's key and payload, and no record is
    // named twice (bench.cu:209-228).
    std::string check() const {
        std::vector<int32_t> got(count_ * words());
        hip_ok(hipMemcpy(got.data(), work_, bytes(), hipMemcpyDeviceToHost), "download");
        std::vector<uint8_t> named(records_ ? count_ : 0, 0);
        for (size_t i = 0; i < count_; ++i) {
            const int32_t k = got[i * words()];
            if (k != sorted_[i]) return "key " + std::to_string(i) + " differs from the host sort";
            if (!records_) continue;
            const int64_t origin = got[i * words() + 1];
            if (origin < 0 || size_t(origin) >= count_) return "payload out of range at " + std::to_string(i);
            if (size_t(origin) / size_t(n_) != i / size_t(n_)) return "record moved to another array";
            if (host_[size_t(origin) * 2] != k) return "record key and payload no longer match";
            if (named[size_t(origin)]++) return "record duplicated";
        }
        return "";
    }

private:
    size_t words() const { return records_ ? 2 : 1; }
    size_t bytes() const { return count_ * words() * sizeof(int32_t); }
    bool records_;
    int n_, arrays_;
    size_t count_;
    const char* dist_;
    std::vector<int32_t> keys_, sorted_, host_;
    void* pristine_ = nullptr;
    void* work_ = nullptr;
};

// ================================================================ jobs
// A configuration to measure: the CSV's descriptive columns plus the launch.
struct Job {
    std::string method;
    int testsorted = -1, fullsm = -1, async = -1;   // -1: not applicable
    int threads = 0, items = 0, working = 0;
    long long shared_bytes = 0, registers = -1, active_blocks = -1, local_bytes = -1;
    std::function<void(const Workload&)> launch;
};

void cs_ok(int rc, const std::string& what) {
    if (rc != CS_OK) fail(what + ": " + cs_strerror(rc));
}

int pow2_at_least(int n) {
    int w = 1;
    while (w < n) w *= 2;
    return w;
}

// Every (method, threads, flags) candidate of the reference's sweep
// (bench.cu:523-552: warp for N <= 32; v1..v6 from N = 32 with threads up to
// min(W, --max-threads)), kept when the library accepts it.
std::vector<Job> bitonic_jobs(const Workload& w, const Settings& s) {
    static const std::pair<int, const char*> methods[] = {
        {CS_BITONIC_WARP, "bitonic_warp"},         {CS_BITONIC_V1, "bitonic_v1"},
        {CS_BITONIC_V2, "bitonic_v2"},             {CS_BITONIC_V3, "bitonic_v3"},
        {CS_BITONIC_V4_XOR, "bitonic_v4_xor"},     {CS_BITONIC_V5_STATIC, "bitonic_v5_static"},
        {CS_BITONIC_V6_REGISTER, "bitonic_v6_register"}};
    const int working = std::max(32, pow2_at_least(w.n()));
    const int top = int(std::min<long long>(working, s.max_threads));
    std::vector<Job> jobs;
    for (const auto& entry : methods) {
        const int method = entry.first;
        const char* name = entry.second;
        if (!s.wants(name)) continue;
        const bool warp = method == CS_BITONIC_WARP;
        if (warp ? w.n() > 32 : w.n() < 32) continue;
        for (int flags = 0; flags < (warp ? 4 : 8); ++flags)
            for (int threads = 32; threads <= (warp ? 32 : top); threads += 32) {
                const uint32_t f = ((flags & 1) ? CS_FLAG_TESTSORTED : 0u) | ((flags & 2) ? CS_FLAG_FULLSM : 0u) |
                                   ((flags & 4) ? CS_FLAG_ASYNC : 0u);
                int32_t bt = 0, ipt = 0, lds = 0, vgprs = -1;
                if (cs_batched_describe(w.dtype(), w.n(), method, threads, f, &bt, &ipt, &lds, &vgprs, nullptr) !=
                    CS_OK)
                    continue;
                Job j;
                j.method = name;
                j.testsorted = flags & 1;
                j.fullsm = (flags >> 1) & 1;
                j.async = warp ? -1 : (flags >> 2) & 1;
                j.threads = threads;
                j.items = ipt;
                j.working = working;
                j.shared_bytes = lds;
                j.registers = vgprs;
                j.launch = [method, threads, f, name](const Workload& wl) {
                    cs_ok(cs_batched_sort(wl.dtype(), wl.data(), wl.data(), wl.arrays(), wl.n(), method, threads, f,
                                          nullptr),
                          name);
                };
                jobs.push_back(std::move(j));
            }
    }
    return jobs;
}

std::vector<Job> baseline_jobs(const Workload& w, const Settings& s, bool block_merge) {
    std::vector<Job> jobs;
    const int type = w.records() ? 1 : 0;
    const int working = std::max(32, pow2_at_least(w.n()));
    if (block_merge && s.wants("cub_block_merge"))
        for (const auto& cfg : csbench::block_merge_configs(type)) {
            if (cfg.threads * cfg.items != working || cfg.threads > s.max_threads) continue;
            csbench::KernelAttrs attrs;
            if (!csbench::block_merge_sort(type, cfg.threads, cfg.items, w.data(), w.n(), 0, nullptr, &attrs)) continue;
            Job j;
            j.method = "cub_block_merge";
            j.threads = cfg.threads;
            j.items = cfg.items;
            j.working = working;
            j.shared_bytes = (long long)attrs.shared_bytes;
            j.registers = attrs.registers;
            j.active_blocks = attrs.active_blocks_per_cu;
            j.local_bytes = attrs.local_bytes;
            j.launch = [type, cfg](const Workload& wl) {
                csbench::block_merge_sort(type, cfg.threads, cfg.items, wl.data(), wl.n(), wl.arrays(), nullptr,
                                          nullptr);
            };
            jobs.push_back(std::move(j));
        }
    if (s.wants("thrust_segment_loop")) {
        Job j;
        j.method = "thrust_segment_loop";
        j.working = w.n();
        j.launch = [type](const Workload& wl) { csbench::thrust_segment_loop(type, wl.data(), wl.n(), wl.arrays(), nullptr); };
        jobs.push_back(std::move(j));
    }
    return jobs;
}

// Single-array sorts: Sort (cs_sort_partition) under every flag combination,
// SortMerge and HybridSort under every TestIfSorted x FullSM pair.
std::vector<Job> large_jobs(const Workload& w, const Settings& s) {
    std::vector<Job> jobs;
    for (int flags = 0; flags < 4; ++flags) {
        const int ts = flags & 1, fs = (flags >> 1) & 1;
        for (int as = 0; as < 2; ++as) {
            if (!s.wants("large_partition")) break;
            const uint32_t f = (ts ? CS_FLAG_TESTSORTED : 0u) | (fs ? CS_FLAG_FULLSM : 0u) | (as ? CS_FLAG_ASYNC : 0u);
            Job j;
            j.method = "large_partition";
            j.testsorted = ts; j.fullsm = fs; j.async = as;
            j.working = w.n();
            j.launch = [f](const Workload& wl) {
                cs_ok(cs_sort_partition(wl.dtype(), wl.data(), wl.n(), 0, 0, f, nullptr), "large_partition");
            };
            jobs.push_back(std::move(j));
        }
        if (s.wants("large_merge")) {
            Job j;
            j.method = "large_merge";
            j.testsorted = ts; j.fullsm = fs;
            j.working = w.n();
            j.launch = [](const Workload& wl) {
                cs_ok(cs_sort_merge(wl.dtype(), wl.data(), wl.n(), 0, 0, nullptr, 0, 1, nullptr), "large_merge");
            };
            jobs.push_back(std::move(j));
        }
        if (s.wants("large_hybrid")) {
            Job j;
            j.method = "large_hybrid";
            j.testsorted = ts; j.fullsm = fs;
            j.working = w.n();
            j.launch = [](const Workload& wl) {
                cs_ok(cs_sort_hybrid(wl.dtype(), wl.data(), wl.n(), nullptr, nullptr), "large_hybrid");
            };
            jobs.push_back(std::move(j));
        }
    }
    return jobs;
}

// ================================================================ CSV
struct Sample {
    const Job& job;
    const Workload& work;
    int trial;
    double event_s, wall_s;
};

struct Column {
    const char* name;
    std::function<void(std::ostream&, const Sample&)> put;
};

double per_comparison(double t, const Sample& s) {
    const double values = double(s.work.n()) * s.work.arrays();
    return s.work.n() > 1 ? t / (values * std::log2(double(s.work.n()))) : std::numeric_limits<double>::quiet_NaN();
}

const std::vector<Column>& columns() {
    static int runtime = -1;
    static const std::vector<Column> table = {
        {"method", [](std::ostream& o, const Sample& s) { o << s.job.method; }},
        {"distribution", [](std::ostream& o, const Sample& s) { o << s.work.dist(); }},
        {"N", [](std::ostream& o, const Sample& s) { o << s.work.n(); }},
        {"arrays", [](std::ostream& o, const Sample& s) { o << s.work.arrays(); }},
        {"total_values", [](std::ostream& o, const Sample& s) { o << (long long)s.work.n() * s.work.arrays(); }},
        {"working_values", [](std::ostream& o, const Sample& s) { o << s.job.working; }},
        {"threads", [](std::ostream& o, const Sample& s) { o << s.job.threads; }},
        {"items_per_thread", [](std::ostream& o, const Sample& s) { o << s.job.items; }},
        {"testsorted", [](std::ostream& o, const Sample& s) { o << s.job.testsorted; }},
        {"fullsm", [](std::ostream& o, const Sample& s) { o << s.job.fullsm; }},
        {"asynccpy", [](std::ostream& o, const Sample& s) { o << s.job.async; }},
        {"shared_bytes", [](std::ostream& o, const Sample& s) { o << s.job.shared_bytes; }},
        {"trial", [](std::ostream& o, const Sample& s) { o << s.trial; }},
        {"E_s", [](std::ostream& o, const Sample& s) { o << s.event_s; }},
        {"X_s", [](std::ostream& o, const Sample& s) { o << per_comparison(s.event_s, s); }},
        {"E_wall_s", [](std::ostream& o, const Sample& s) { o << s.wall_s; }},
        {"wall_X_s", [](std::ostream& o, const Sample& s) { o << per_comparison(s.wall_s, s); }},
        {"time_per_array_s", [](std::ostream& o, const Sample& s) { o << s.event_s / s.work.arrays(); }},
        {"time_per_value_s",
         [](std::ostream& o, const Sample& s) { o << s.event_s / (double(s.work.n()) * s.work.arrays()); }},
        {"validated", [](std::ostream& o, const Sample&) { o << 1; }},
        {"cuda_runtime",
         [](std::ostream& o, const Sample&) {
             if (runtime < 0) hip_ok(hipRuntimeGetVersion(&runtime), "hipRuntimeGetVersion");
             o << runtime;
         }},
        {"thrust_version", [](std::ostream& o, const Sample&) { o << csbench::thrust_version(); }},
        {"cub_version", [](std::ostream& o, const Sample&) { o << csbench::hipcub_version(); }},
        {"registers_per_thread", [](std::ostream& o, const Sample& s) { o << s.job.registers; }},
        {"active_blocks_per_sm", [](std::ostream& o, const Sample& s) { o << s.job.active_blocks; }},
        {"local_bytes_per_thread", [](std::ostream& o, const Sample& s) { o << s.job.local_bytes; }},
    };
    return table;
}

class CsvSink {
public:
    explicit CsvSink(const std::string& path) {
        const auto dir = std::filesystem::path(path).parent_path();
        if (!dir.empty()) std::filesystem::create_directories(dir);
        file_.open(path);
        if (!file_) fail("Cannot open " + path);
        file_.precision(17);
        const auto& cols = columns();
        for (size_t c = 0; c < cols.size(); ++c) file_ << (c ? "," : "") << cols[c].name;
        file_ << '\n';
    }
    void add(const Sample& s) {
        const auto& cols = columns();
        for (size_t c = 0; c < cols.size(); ++c) {
            if (c) file_ << ',';
            cols[c].put(file_, s);
        }
        file_ << '\n';
        file_.flush();
    }

private:
    std::ofstream file_;
};

// ================================================================ running
// Warm-up launch and full check outside the timers; then each timed trial
// starts from the restored input and its own output is checked afterwards.
void measure(const Job& job, const Workload& w, const Settings& s, CsvSink& sink) {
    auto verified = [&] {
        const std::string problem = w.check();
        if (!problem.empty())
            fail("Validation failed: " + job.method + " N=" + std::to_string(w.n()) + " threads=" +
                 std::to_string(job.threads) + " sorted=" + std::to_string(job.testsorted) + " fullsm=" +
                 std::to_string(job.fullsm) + " async=" + std::to_string(job.async) + " distribution=" + w.dist() +
                 " (" + problem + ")");
    };
    w.restore();
    job.launch(w);
    hip_ok(hipGetLastError(), "launch");
    hip_ok(hipDeviceSynchronize(), "launch");
    verified();
    if (s.check_only) {
        std::cout << "PASS " << job.method << " N=" << w.n() << " threads=" << job.threads
                  << " sorted=" << job.testsorted << " fullsm=" << job.fullsm << " async=" << job.async << ' '
                  << w.dist() << '\n';
        return;
    }
    hipEvent_t start, stop;
    hip_ok(hipEventCreate(&start), "event");
    hip_ok(hipEventCreate(&stop), "event");
    for (int trial = 0; trial < int(s.repetitions); ++trial) {
        w.restore();
        const auto t0 = std::chrono::steady_clock::now();
        hip_ok(hipEventRecord(start, nullptr), "event");
        job.launch(w);
        hip_ok(hipGetLastError(), "launch");
        hip_ok(hipEventRecord(stop, nullptr), "event");
        hip_ok(hipEventSynchronize(stop), "event");
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        float ms = 0.f;
        hip_ok(hipEventElapsedTime(&ms, start, stop), "event");
        verified();
        sink.add(Sample{job, w, trial, double(ms) * 1e-3, wall});
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
}

std::vector<const Distribution*> active_distributions(const Settings& s) {
    std::vector<const Distribution*> out;
    for (const auto& d : distributions())
        if (!d.check_only || s.check_only) out.push_back(&d);
    return out;
}

// The reference's length sweep (bench.cu:587-598): 1..32, multiples of 32 up
// to 1024, every power of two from 32 and its neighbours, and the maximum.
std::vector<int> sweep_lengths(int lo, int hi) {
    std::vector<int> v;
    auto add = [&](long long n) {
        if (n >= 1 && n <= hi) v.push_back(int(n));
    };
    for (int n = 1; n <= 32; ++n) add(n);
    for (int n = 64; n <= 1024; n += 32) add(n);
    for (long long p = 32; p <= hi; p *= 2) {
        add(p - 1);
        add(p);
        add(p + 1);
    }
    add(hi);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    v.erase(v.begin(), std::lower_bound(v.begin(), v.end(), lo));
    return v;
}

// Arrays per batch (bench.cu:605-608): at most --arrays, at most
// --batch-elements values, and an eighth of free device memory.
int batch_arrays(int n, size_t elem_bytes, const Settings& s) {
    size_t free_b = 0, total_b = 0;
    hip_ok(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t values = std::min(size_t(s.batch_elements), free_b / (8 * elem_bytes));
    if (values < size_t(n)) fail("Batch/memory budget cannot hold one array; increase --batch-elements");
    return int(std::min(size_t(s.arrays), values / size_t(n)));
}

void run_small(const Settings& s, bool records, const std::string& tname) {
    const int capacity = cs_batched_max_n(records ? CS_REC_I32_I32 : CS_I32);
    if (s.mode == "small" && s.max_size > capacity)
        fail("--max-size exceeds the small bitonic capacity for " + tname + " (" + std::to_string(capacity) + ")");
    const int hi = s.max_size ? int(std::min<long long>(s.max_size, capacity)) : capacity;
    if (s.min_size > hi) return;
    CsvSink sink(s.prefix + "_" + tname + "_small.csv");
    for (const int n : sweep_lengths(int(s.min_size), hi)) {
        const int arrays = batch_arrays(n, records ? 8 : 4, s);
        for (const Distribution* d : active_distributions(s)) {
            std::cout << "small " << tname << " N=" << n << " arrays=" << arrays << ' ' << d->name << std::endl;
            const Workload w(records, n, arrays, *d);
            for (const auto& job : bitonic_jobs(w, s)) measure(job, w, s, sink);
            for (const auto& job : baseline_jobs(w, s, true)) measure(job, w, s, sink);
        }
    }
}

void run_large(const Settings& s, bool records, const std::string& tname) {
    const size_t elem = records ? 8 : 4;
    size_t free_b = 0, total_b = 0;
    hip_ok(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const long long by_memory = std::min<long long>(1LL << 28, (long long)(free_b / (8 * elem)));
    const long long hi = std::min<long long>(s.max_size ? s.max_size : (1LL << 24), by_memory);
    const long long lo = std::max<long long>(s.min_size, cs_batched_max_n(records ? CS_REC_I32_I32 : CS_I32) + 1);
    CsvSink sink(s.prefix + "_" + tname + "_large.csv");
    for (long long n = pow2_at_least(int(lo)); n <= hi; n *= 2)
        for (const Distribution* d : active_distributions(s)) {
            std::cout << "large " << tname << " N=" << n << ' ' << d->name << std::endl;
            const Workload w(records, int(n), 1, *d);
            for (const auto& job : large_jobs(w, s)) measure(job, w, s, sink);
            for (const auto& job : baseline_jobs(w, s, false)) measure(job, w, s, sink);
        }
}

}  // namespace bench

int main(int argc, char** argv) {
    try {
        const bench::Settings s = bench::parse_command_line(argc, argv);
        bench::hip_ok(hipSetDevice(0), "hipSetDevice");
        for (const bool records : {false, true}) {
            const std::string tname = records ? "record" : "int";
            if (s.type != "both" && s.type != tname) continue;
            if (s.mode != "large") bench::run_small(s, records, tname);
            if (s.mode != "small") bench::run_large(s, records, tname);
        }
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
