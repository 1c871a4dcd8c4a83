// Uses the C++ drop-in header (cuda-sort_amd/cusort.hpp) the way the
// reference's callers use cusort.hpp (test.cu:58-61, 471, 525, 578; bench.cu
// 453-534, 660-666): device buffers, the template entry points, host
// std::sort as the check.  Exit status 0 when every case matches.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../cuda-sort_amd/cusort.hpp"

using cusort_hip::IntAndData;

// A caller's own crossover, as the reference invites (cusort.hpp:2476-2478):
// HybridSort must consult it.
template <>
struct cusort_hip::CuSortTuning<unsigned int> {
    static constexpr int PartitionThreshold = 1 << 16;
};
// The default trait is a compile-time constant, usable as the reference's
// dispatcher uses it (cusort.hpp:2667-2668).
static_assert(cusort_hip::CuSortTuning<int>::PartitionThreshold > (1 << 20), "int threshold is a constant");
static_assert(cusort_hip::CuSortTuning<unsigned int>::PartitionThreshold == (1 << 16), "specialised");

static int failures = 0;

template <class T> bool key_less(const T& a, const T& b) { return a < b; }

template <class T>
void expect_sorted_equal(std::vector<T> in, const std::vector<T>& got, int n_per, const char* what) {
    for (size_t a = 0; a < in.size(); a += size_t(n_per))
        std::sort(in.begin() + long(a), in.begin() + long(std::min(in.size(), a + size_t(n_per))), key_less<T>);
    for (size_t i = 0; i < in.size(); ++i) {
        const bool same = !(in[i] < got[i]) && !(got[i] < in[i]);
        if (!same) {
            std::printf("FAIL %s at %zu\n", what, i);
            ++failures;
            return;
        }
    }
    std::printf("PASS %s\n", what);
}

template <class T> T* upload(const std::vector<T>& h) {
    T* d = nullptr;
    if (hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(T)) != hipSuccess) std::exit(2);
    if (!h.empty()) (void)hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    return d;
}
template <class T> std::vector<T> download(const T* d, size_t n) {
    std::vector<T> h(n);
    (void)hipDeviceSynchronize();
    if (n) (void)hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost);
    return h;
}

int main() {
    std::mt19937 rng(1234);
    // batched: one array per workgroup, packed (bench.cu:355-393)
    for (int n : {7, 32, 100, 1024, 4096}) {
        const int arrays = 37;
        std::vector<int> h(static_cast<size_t>(n) * arrays);
        for (auto& v : h) v = int(rng());
        int* d = upload(h);
        if (n <= 32) cusort_hip::Sort_1Warp_SM_le32_batched<int>(d, n, arrays);
        else cusort_hip::Sort_1Block_batched<int, 3, true, false, false>(d, n, arrays, 64);
        expect_sorted_equal(h, download(d, h.size()), n, "batched v3/warp int");
        (void)hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
        if (n == 4096) {   // Working must be max(32, pow2(n)) (bench.cu:497-521)
            cusort_hip::Sort_1Block_v6_register_batched<int, 4096, 256>(d, n, arrays);
            expect_sorted_equal(h, download(d, h.size()), n, "batched v6 int");
        }
        (void)hipFree(d);
    }
    {
        const int n = 1000, arrays = 50;
        std::vector<IntAndData> h(static_cast<size_t>(n) * arrays);
        for (size_t i = 0; i < h.size(); ++i) h[i] = IntAndData{int(rng() % 100), int(i)};
        IntAndData* d = upload(h);
        cusort_hip::Sort_1Block_v5_static_batched<IntAndData, 1024, 128>(d, n, arrays);
        expect_sorted_equal(h, download(d, h.size()), n, "batched v5 IntAndData");
        (void)hipFree(d);
    }
    // large arrays (test.cu:444-602 sizes)
    for (int n : {33, 5000, 100000, 1 << 20}) {
        std::vector<int> h(static_cast<size_t>(n));
        for (auto& v : h) v = int(rng() % 1000000);
        int* d = upload(h);
        cusort_hip::SortMerge<int>(d, n);
        expect_sorted_equal(h, download(d, h.size()), n, "SortMerge int");
        (void)hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
        cusort_hip::Sort<int>(d, n);
        expect_sorted_equal(h, download(d, h.size()), n, "Sort int");
        (void)hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
        cusort_hip::CuSortWorkspace<int> ws{static_cast<size_t>(n)};
        cusort_hip::HybridSort<int>(d, n, ws);
        expect_sorted_equal(h, download(d, h.size()), n, "HybridSort int (workspace)");
        (void)hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
        cusort_hip::HybridSortV3<int>(d, n);
        expect_sorted_equal(h, download(d, h.size()), n, "HybridSortV3 int");
        int* out = upload(std::vector<int>(size_t(n)));
        (void)hipMemcpy(d, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
        cusort_hip::SortMerge<int>(d, n, 0, nullptr, out, size_t(n), false);
        expect_sorted_equal(h, download(out, h.size()), n, "SortMerge out of place");
        (void)hipFree(out);
        (void)hipFree(d);
    }
    {
        const int n = 300000;
        std::vector<unsigned long long> h(static_cast<size_t>(n));
        for (auto& v : h) v = (uint64_t(rng()) << 32) | rng();
        unsigned long long* d = upload(h);
        cusort_hip::CuSortWorkspace<unsigned long long> ws{static_cast<size_t>(n)};
        cusort_hip::HybridSortPartitionedV3<unsigned long long>(d, n, ws);
        expect_sorted_equal(h, download(d, h.size()), n, "HybridSortPartitionedV3 u64");
        (void)hipFree(d);
    }
    {   // the tuning trait decides the path; the workspace path allocates nothing
        const int n = 1 << 18;
        std::vector<unsigned int> h(static_cast<size_t>(n));
        for (auto& v : h) v = rng();
        unsigned int* d = upload(h);
        cusort_hip::CuSortWorkspace<unsigned int> ws{static_cast<size_t>(n)};
        int64_t merges0 = 0, parts0 = 0, merges1 = 0, parts1 = 0, allocs0 = 0, allocs1 = 0;
        cs_workspace_calls(ws.handle(), &merges0, &parts0);
        cs_memory_stats(&allocs0, nullptr);
        if (n >= cusort_hip::CuSortTuning<unsigned int>::PartitionThreshold)   // reference-style use of the trait
            cusort_hip::HybridSort<unsigned int>(d, n, ws);
        cs_workspace_calls(ws.handle(), &merges1, &parts1);
        cs_memory_stats(&allocs1, nullptr);
        expect_sorted_equal(h, download(d, h.size()), n, "HybridSort with a specialised CuSortTuning");
        const bool partitioned = parts1 == parts0 + 1 && merges1 == merges0;
        std::printf("%s specialised threshold takes the partitioned path\n", partitioned ? "PASS" : "FAIL");
        std::printf("%s workspace-backed partition sort allocates nothing (%lld allocations)\n",
                    allocs1 == allocs0 ? "PASS" : "FAIL", (long long)(allocs1 - allocs0));
        if (!partitioned || allocs1 != allocs0) ++failures;
        // the same n as int stays below int's threshold: merge path, no allocation either
        std::vector<int> hi(static_cast<size_t>(n));
        for (auto& v : hi) v = int(rng());
        int* di = upload(hi);
        cusort_hip::CuSortWorkspace<int> wsi{static_cast<size_t>(n)};
        cs_memory_stats(&allocs0, nullptr);
        cusort_hip::HybridSortV3<int>(di, n, wsi);
        cs_memory_stats(&allocs1, nullptr);
        expect_sorted_equal(hi, download(di, hi.size()), n, "HybridSortV3 int (workspace, merge path)");
        int64_t m = 0, p = 0;
        cs_workspace_calls(wsi.handle(), &m, &p);
        const bool merged = m == 1 && p == 0 && allocs1 == allocs0;
        std::printf("%s default threshold keeps the merge path without allocating\n", merged ? "PASS" : "FAIL");
        if (!merged) ++failures;
        // reference accessors
        bool acc = ws.valuesBuffer() != nullptr && ws.capacity() >= size_t(n) && ws.nbSortStreams() == 8 &&
                   ws.persistentBlocks() > 0 && ws.pivots() != nullptr && ws.counters() != nullptr &&
                   ws.blockOffsets() != nullptr && ws.workerEventsRecorded();
        hipStream_t worker = ws.sortStream(3);
        acc = acc && worker != nullptr && hipStreamWaitEvent(worker, ws.sortDone(0), 0) == hipSuccess &&
              hipStreamSynchronize(worker) == hipSuccess;
        ws.setBlocksPerSM(2);
        acc = acc && ws.persistentBlocks() > 0 && ws.blockOffsets() != nullptr;
        std::printf("%s workspace accessors\n", acc ? "PASS" : "FAIL");
        if (!acc) ++failures;
        (void)hipFree(d);
        (void)hipFree(di);
    }
    {   // key / payload pairs in two arrays: each pair must survive, keys in order
        const int n = 1000, arrays = 40;
        std::vector<long long> hk(static_cast<size_t>(n) * arrays);
        std::vector<int> hv(hk.size());
        for (size_t i = 0; i < hk.size(); ++i) {
            hk[i] = static_cast<long long>(rng() % 50) - 25;
            hv[i] = int(i);
        }
        long long* dk = upload(hk);
        int* dv = upload(hv);
        cusort_hip::SortPairs_batched<long long>(dk, dv, n, arrays);
        const std::vector<long long> gk = download(dk, hk.size());
        const std::vector<int> gv = download(dv, hv.size());
        bool good = true;
        for (size_t i = 0; i < gk.size() && good; ++i) {
            good = gv[i] >= 0 && size_t(gv[i]) < hk.size() && hk[size_t(gv[i])] == gk[i] &&
                   size_t(gv[i]) / size_t(n) == i / size_t(n);
            if (good && i % size_t(n) != 0) good = gk[i - 1] <= gk[i];
        }
        std::printf("%s SortPairs_batched int64\n", good ? "PASS" : "FAIL");
        if (!good) ++failures;
        (void)hipFree(dk);
        (void)hipFree(dv);
    }
    {   // fp64 keys (finite values: operator< agrees with totalOrder up to -0 == +0)
        std::normal_distribution<double> normal(0.0, 1e3);
        for (int n : {1000, 200000}) {
            std::vector<double> h(static_cast<size_t>(n));
            for (auto& v : h) v = normal(rng);
            double* d = upload(h);
            cusort_hip::Sort<double>(d, n);
            expect_sorted_equal(h, download(d, h.size()), n, "Sort double");
            (void)hipFree(d);
        }
    }
    std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
