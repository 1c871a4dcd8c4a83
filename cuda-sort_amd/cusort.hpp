// cusort-hip C++ drop-in layer: the reference's entry-point names over the
// C ABI of libcusort_hip.so (include/cusort_hip.h).
//
// Reference interface (berenger-eu/cuda-sort)        here
//   IntAndData (intanddata.hpp:9-79)                 cusort_hip::IntAndData (same 8-byte layout, key order)
//   Sort<T,...>(T*, int, stream) (cusort.hpp:2020)   Sort<T,...>            -> cs_sort_partition
//   SortMerge<T,...>(...) (cusort.hpp:2161-2170)      SortMerge<T,...>       -> cs_sort_merge
//   HybridSort / HybridSortV3 (2656-2754)            HybridSort / HybridSortV3 -> cs_sort_hybrid
//   HybridSortPartitionedV3 (2488-2497)              HybridSortPartitionedV3 -> cs_sort_partition
//   HybridSortLegacy (2355-2356, never called)       HybridSortLegacy       -> cs_sort_hybrid
//   CuSortWorkspace<T,MaxNbParts> (67-219)           CuSortWorkspace<T,MaxNbParts> (RAII, move-only)
//   CuSortTuning<T> (2479-2483)                      CuSortTuning<T>        (specialisable constexpr threshold)
//   batched kernels of bench.cu:355-393, launched    Sort_1Warp_SM_le32_batched, Sort_1Block_batched<Variant>,
//   kernel<<<arrays, threads, smem>>>(data, n)        Sort_1Block_v4_xor_batched, Sort_1Block_v5_static_batched,
//                                                    Sort_1Block_v6_register_batched -> cs_batched_sort
//
// Compile-time switches of the reference (TestIfSorted, FullSM, Async) are
// kept as template parameters and passed to the library as run-time flags.
// Errors: like the reference's CUDA_CHECK (cusort.hpp:240-245) a failing call
// prints the status and exits; define CUSORT_HIP_THROW before including this
// header to get a std::runtime_error instead.
#ifndef CUSORT_HIP_CUSORT_HPP
#define CUSORT_HIP_CUSORT_HPP

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../include/cusort_hip.h"

namespace cusort_hip {

struct IntAndData {
    int val;
    int data;
    IntAndData& operator=(const int& v) { val = v; data = 0; return *this; }
    bool operator<(const IntAndData& o) const { return val < o.val; }
    bool operator<=(const IntAndData& o) const { return val <= o.val; }
    bool operator>(const IntAndData& o) const { return val > o.val; }
    bool operator>=(const IntAndData& o) const { return val >= o.val; }
    bool operator==(const IntAndData& o) const { return val == o.val && data == o.data; }
    bool operator!=(const IntAndData& o) const { return !(*this == o); }
};
static_assert(sizeof(IntAndData) == 8, "IntAndData must keep the reference's 8-byte layout");

// {int64 key, int32 payload} record of the serving configuration (16 bytes).
struct Int64AndData {
    int64_t key;
    int32_t data;
    int32_t pad;
};
static_assert(sizeof(Int64AndData) == 16, "16-byte record layout");

template <class T> struct DType;
template <> struct DType<int> { static constexpr int value = CS_I32; };
template <> struct DType<unsigned int> { static constexpr int value = CS_U32; };
template <> struct DType<float> { static constexpr int value = CS_F32; };
template <> struct DType<double> { static constexpr int value = CS_F64; };
template <> struct DType<long> { static constexpr int value = sizeof(long) == 8 ? CS_I64 : CS_I32; };
template <> struct DType<long long> { static constexpr int value = CS_I64; };
template <> struct DType<unsigned long> { static constexpr int value = sizeof(unsigned long) == 8 ? CS_U64 : CS_U32; };
template <> struct DType<unsigned long long> { static constexpr int value = CS_U64; };
template <> struct DType<IntAndData> { static constexpr int value = CS_REC_I32_I32; };
template <> struct DType<Int64AndData> { static constexpr int value = CS_REC_I64_I32; };

inline void check(int status, const char* what) {
    if (status == CS_OK) return;
#ifdef CUSORT_HIP_THROW
    throw std::runtime_error(std::string(what) + ": " + cs_strerror(status));
#else
    std::fprintf(stderr, "cusort-hip error in %s: %s (%d)\n", what, cs_strerror(status), status);
    std::exit(EXIT_FAILURE);
#endif
}

// W = max(32, next power of two >= n), the working length every batched
// kernel of the reference is instantiated for (bench.cu:524).
inline int working_length(int n) {
    int w = 32;
    while (w < n) w *= 2;
    return w;
}

template <bool TestIfSorted, bool FullSM, bool Async>
constexpr uint32_t flags() {
    return (TestIfSorted ? CS_FLAG_TESTSORTED : 0u) | (FullSM ? CS_FLAG_FULLSM : 0u) | (Async ? CS_FLAG_ASYNC : 0u);
}

// ---------------------------------------------------------- workspace ----
// CuSortWorkspace<T, MaxNbParts> (reference cusort.hpp:67-219): RAII and
// move-only over a cs_workspace.  The library keeps the values buffer, an
// arena for every other temporary of a sort of up to capacity() elements and
// nbSortStreams() worker streams with completion events; sorts through a
// workspace allocate no device memory.  The reference's scratch accessors
// (pivots, counters, blockOffsets, ...) return regions of the arena sized as
// the reference's arrays (MaxNbParts entries, persistentBlocks x MaxNbParts
// for the per-block ones); the library's kernels keep their own data there,
// so their contents are not the reference's.
template <typename T, int MaxNbParts = 1024>
class CuSortWorkspace {
    static_assert(MaxNbParts >= 32 && MaxNbParts <= 1024 && ((MaxNbParts - 1) & MaxNbParts) == 0,
                  "MaxNbParts must be a power of two in [32, 1024]");
    cs_workspace* ws_ = nullptr;
    std::vector<int> host_counters_ = std::vector<int>(MaxNbParts + 1);
    std::vector<int> host_uniform_parts_ = std::vector<int>(MaxNbParts);
    std::vector<int> host_large_partitions_ = std::vector<int>(MaxNbParts);

    template <class P>
    P* arena_at(size_t byte_offset) const {
        void* arena = nullptr;
        size_t bytes = 0;
        cs_workspace_info(ws_, nullptr, &arena, &bytes, nullptr, nullptr);
        return arena && byte_offset < bytes ? reinterpret_cast<P*>(static_cast<char*>(arena) + byte_offset) : nullptr;
    }
    // Byte offsets of the reference-shaped regions inside the arena.
    size_t off_counters() const { return size_t(MaxNbParts) * sizeof(T); }
    size_t off_block_offsets() const { return off_counters() + size_t(MaxNbParts + 2) * sizeof(int); }
    size_t off_block_uniform() const {
        return off_block_offsets() + size_t(persistentBlocks()) * MaxNbParts * sizeof(int);
    }
    size_t off_uniform_parts() const {
        return off_block_uniform() + size_t(persistentBlocks()) * MaxNbParts * sizeof(int);
    }
    size_t reference_regions() const { return off_uniform_parts() + size_t(MaxNbParts + 1) * sizeof(int); }
    void keep_regions() {
        check(cs_workspace_reserve_scratch(ws_, reference_regions()), "CuSortWorkspace (scratch regions)");
    }

public:
    explicit CuSortWorkspace(size_t initialCapacity = 0, int blocksPerSM = 4, int nbSortStreams = 8) {
        check(cs_workspace_create(&ws_, DType<T>::value, initialCapacity, blocksPerSM < 1 ? 1 : blocksPerSM,
                                  nbSortStreams < 1 ? 1 : nbSortStreams),
              "CuSortWorkspace");
        keep_regions();
    }
    ~CuSortWorkspace() {
        if (ws_) cs_workspace_destroy(ws_);
    }
    CuSortWorkspace(const CuSortWorkspace&) = delete;
    CuSortWorkspace& operator=(const CuSortWorkspace&) = delete;
    CuSortWorkspace(CuSortWorkspace&& o) noexcept
        : ws_(o.ws_), host_counters_(std::move(o.host_counters_)),
          host_uniform_parts_(std::move(o.host_uniform_parts_)),
          host_large_partitions_(std::move(o.host_large_partitions_)) {
        o.ws_ = nullptr;
    }

    void reserve(size_t capacity) {
        check(cs_workspace_reserve(ws_, capacity), "CuSortWorkspace::reserve");
        keep_regions();
    }
    void setBlocksPerSM(int blocksPerSM) {
        check(cs_workspace_set_blocks_per_cu(ws_, blocksPerSM < 1 ? 1 : blocksPerSM), "CuSortWorkspace::setBlocksPerSM");
        keep_regions();
    }

    T* valuesBuffer() {
        void* v = nullptr;
        cs_workspace_info(ws_, &v, nullptr, nullptr, nullptr, nullptr);
        return static_cast<T*>(v);
    }
    T* pivots() { return arena_at<T>(0); }
    int* counters() { return arena_at<int>(off_counters()); }
    int* skipScatterFlag() { return counters() ? counters() + MaxNbParts + 1 : nullptr; }
    int* blockOffsets() { return arena_at<int>(off_block_offsets()); }
    int* blockUniform() { return arena_at<int>(off_block_uniform()); }
    int* uniformParts() { return arena_at<int>(off_uniform_parts()); }
    int* activeNbParts() { return arena_at<int>(off_uniform_parts() + size_t(MaxNbParts) * sizeof(int)); }
    size_t capacity() const {
        size_t c = 0;
        cs_workspace_capacity(ws_, &c);
        return c;
    }
    int persistentBlocks() const {
        int32_t b = 0;
        cs_workspace_info(ws_, nullptr, nullptr, nullptr, &b, nullptr);
        return b;
    }
    std::vector<int>& hostCounters() { return host_counters_; }
    std::vector<int>& hostUniformParts() { return host_uniform_parts_; }
    std::vector<int>& hostLargePartitions() { return host_large_partitions_; }
    int nbSortStreams() const {
        int32_t n = 0;
        cs_workspace_info(ws_, nullptr, nullptr, nullptr, nullptr, &n);
        return n;
    }
    hipStream_t sortStream(int idx) const {
        hipStream_t st = nullptr;
        check(cs_workspace_stream(ws_, idx, &st, nullptr), "CuSortWorkspace::sortStream");
        return st;
    }
    hipEvent_t sortDone(int idx) const {
        cs_hip_event ev = nullptr;
        check(cs_workspace_stream(ws_, idx, nullptr, &ev), "CuSortWorkspace::sortDone");
        return ev;
    }
    bool workerEventsRecorded() const {
        int32_t r = 0;
        cs_workspace_events_recorded(ws_, &r);
        return r != 0;
    }
    // The library records the events itself; kept for source compatibility.
    void markWorkerEventsRecorded() {}
    cs_workspace* handle() { return ws_; }
};

// CuSortTuning<T> (reference cusort.hpp:2476-2483): the element count from
// which HybridSort / HybridSortV3 take the partitioned path, a specialisable
// compile-time constant.  The default is the crossover measured on MI355X
// (merge below, partition above: 2^23 for 4- and 8-byte elements, 2^21 for
// 16-byte records), where the reference's default (tuned on its own GPUs)
// never partitions 4-byte keys and partitions 8-byte ones from 2^27.  A
// caller's specialisation is what HybridSort consults.
template <typename T>
struct CuSortTuning {
    static constexpr int PartitionThreshold = sizeof(T) >= 16 ? (1 << 21) : (1 << 23);
};

// ------------------------------------------------------- large arrays ----
template <typename T, bool TestIfSorted = false, bool FullSM = false, bool ASyncCpyBito = false,
          int MinNbParts = 32, int MaxNbParts = 1024>
void Sort(T* dataToSort, const int64_t nbValuesToSort, hipStream_t stream = nullptr) {
    check(cs_sort_partition(DType<T>::value, dataToSort, nbValuesToSort, MinNbParts, MaxNbParts,
                            flags<TestIfSorted, FullSM, ASyncCpyBito>(), stream),
          "Sort");
}

template <typename T, int MaxValuesToSort = 2048, bool TestIfSorted = false, bool FullSM = false>
void SortMerge(T* dataToSort, const int64_t nbValuesToSort, const int nbThreadsPerBlockForMerge = 0,
               hipStream_t stream = nullptr, T* externalBuffer = nullptr, size_t externalBufferElems = 0,
               bool storedInRightPlaced = true) {
    static_assert(((MaxValuesToSort - 1) & MaxValuesToSort) == 0, "MaxValuesToSort must be a power of 2");
    check(cs_sort_merge(DType<T>::value, dataToSort, nbValuesToSort, MaxValuesToSort, nbThreadsPerBlockForMerge,
                        externalBuffer, externalBufferElems, storedInRightPlaced ? 1 : 0, stream),
          "SortMerge");
}

// Explicit partitioned path with the workspace's buffers (no allocation).
template <typename T, int MaxValuesToSort = 2048, int MaxNbParts = 1024, bool TestIfSorted = false,
          bool FullSM = true>
void HybridSortPartitionedV3(T* dataToSort, const int64_t nbValuesToSort, CuSortWorkspace<T, MaxNbParts>& workspace,
                             hipStream_t stream = nullptr) {
    static_assert(((MaxValuesToSort - 1) & MaxValuesToSort) == 0, "MaxValuesToSort must be a power of 2");
    if (nbValuesToSort <= 1) return;
    workspace.reserve(size_t(nbValuesToSort));
    check(cs_sort_partition_ws(DType<T>::value, dataToSort, nbValuesToSort, 0, MaxNbParts,
                               flags<TestIfSorted, FullSM, false>(), workspace.handle(), stream),
          "HybridSortPartitionedV3");
}

// Dispatch as the reference's HybridSortV3 (cusort.hpp:2656-2703): below
// CuSortTuning<T>::PartitionThreshold the merge sort (single tiles need no
// workspace), above it the partitioned path; the workspace grows to n.
template <typename T, int MaxValuesToSort = 2048, int MaxNbParts = 1024, bool TestIfSorted = false,
          bool FullSM = true>
void HybridSortV3(T* dataToSort, const int64_t nbValuesToSort, CuSortWorkspace<T, MaxNbParts>& workspace,
                  hipStream_t stream = nullptr) {
    if (nbValuesToSort <= 1) return;
    if (nbValuesToSort >= CuSortTuning<T>::PartitionThreshold) {
        HybridSortPartitionedV3<T, MaxValuesToSort, MaxNbParts, TestIfSorted, FullSM>(dataToSort, nbValuesToSort,
                                                                                     workspace, stream);
        return;
    }
    if (nbValuesToSort <= cs_batched_max_n(DType<T>::value)) {
        SortMerge<T, MaxValuesToSort, TestIfSorted, FullSM>(dataToSort, nbValuesToSort, 0, stream);
        return;
    }
    workspace.reserve(size_t(nbValuesToSort));
    check(cs_sort_merge_ws(DType<T>::value, dataToSort, nbValuesToSort, workspace.handle(), stream), "HybridSortV3");
}

// Without a workspace (reference cusort.hpp:2706-2725): the partitioned path
// builds a temporary workspace and synchronises the stream before it goes.
template <typename T, int MaxValuesToSort = 2048, int MaxNbParts = 1024, bool TestIfSorted = false,
          bool FullSM = true>
void HybridSortV3(T* dataToSort, const int64_t nbValuesToSort, hipStream_t stream = nullptr) {
    if (nbValuesToSort < CuSortTuning<T>::PartitionThreshold) {
        SortMerge<T, MaxValuesToSort, TestIfSorted, FullSM>(dataToSort, nbValuesToSort, 0, stream);
        return;
    }
    CuSortWorkspace<T, MaxNbParts> workspace{static_cast<size_t>(nbValuesToSort)};
    HybridSortPartitionedV3<T, MaxValuesToSort, MaxNbParts, TestIfSorted, FullSM>(dataToSort, nbValuesToSort,
                                                                                 workspace, stream);
    check(int(hipStreamSynchronize(stream)), "HybridSortV3");
}

template <typename T, int MaxValuesToSort = 2048, int MaxNbParts = 1024, bool TestIfSorted = false,
          bool FullSM = true>
void HybridSort(T* dataToSort, const int64_t nbValuesToSort, CuSortWorkspace<T, MaxNbParts>& workspace,
                hipStream_t stream = nullptr) {
    HybridSortV3<T, MaxValuesToSort, MaxNbParts, TestIfSorted, FullSM>(dataToSort, nbValuesToSort, workspace,
                                                                        stream);
}

template <typename T, int MaxValuesToSort = 2048, int MaxNbParts = 1024, bool TestIfSorted = false,
          bool FullSM = true>
void HybridSort(T* dataToSort, const int64_t nbValuesToSort, hipStream_t stream = nullptr) {
    HybridSortV3<T, MaxValuesToSort, MaxNbParts, TestIfSorted, FullSM>(dataToSort, nbValuesToSort, stream);
}

template <typename T, int MaxValuesToSort = 2048, int MaxNbParts = 512, bool TestIfSorted = false,
          bool FullSM = false>
void HybridSortLegacy(T* dataToSort, const int64_t nbValuesToSort, hipStream_t stream = nullptr,
                      T* dataBuffer = nullptr) {
    (void)dataBuffer;
    check(cs_sort_hybrid(DType<T>::value, dataToSort, nbValuesToSort, nullptr, stream), "HybridSortLegacy");
}

// ------------------------------------------------------------ batched ----
// The reference launches these as kernel<<<arrays, threads, smem>>>(data, n)
// on `arrays` packed arrays of n elements (bench.cu:355-393, 442-552).
template <typename T, bool Sorted = false, bool Shared = false>
void Sort_1Warp_SM_le32_batched(T* data, int n, int64_t arrays, hipStream_t stream = nullptr) {
    check(cs_batched_sort(DType<T>::value, data, data, arrays, n, CS_BITONIC_WARP, 32,
                          flags<Sorted, Shared, false>(), stream),
          "Sort_1Warp_SM_le32_batched");
}

// Variant 1, 2 or 3 of Sort_1Block_[async_]gt32_d_extsm[_v2|_v3].
template <typename T, int Variant, bool Sorted = false, bool Shared = false, bool Async = false>
void Sort_1Block_batched(T* data, int n, int64_t arrays, int threads = 0, hipStream_t stream = nullptr) {
    static_assert(Variant >= 1 && Variant <= 3, "variant 1, 2 or 3");
    check(cs_batched_sort(DType<T>::value, data, data, arrays, n, CS_BITONIC_V1 + (Variant - 1), threads,
                          flags<Sorted, Shared, Async>(), stream),
          "Sort_1Block_batched");
}

template <typename T, bool Sorted = false, bool Shared = false, bool Async = false>
void Sort_1Block_v4_xor_batched(T* data, int n, int64_t arrays, int threads, hipStream_t stream = nullptr) {
    check(cs_batched_sort(DType<T>::value, data, data, arrays, n, CS_BITONIC_V4_XOR, threads,
                          flags<Sorted, Shared, Async>(), stream),
          "Sort_1Block_v4_xor_batched");
}

template <typename T, unsigned Working, unsigned Threads, bool Sorted = false, bool Shared = false,
          bool Async = false>
void Sort_1Block_v5_static_batched(T* data, int n, int64_t arrays, hipStream_t stream = nullptr) {
    static_assert(Working >= Threads && Threads >= 32 && (Threads & (Threads - 1)) == 0, "invalid static dimensions");
    if (int(Working) != working_length(n)) check(CS_E_UNSUPPORTED, "Sort_1Block_v5_static_batched (Working != W(n))");
    check(cs_batched_sort(DType<T>::value, data, data, arrays, n, CS_BITONIC_V5_STATIC, int(Threads),
                          flags<Sorted, Shared, Async>(), stream),
          "Sort_1Block_v5_static_batched");
}

template <typename T, unsigned Working, unsigned Threads, bool Sorted = false, bool Shared = false,
          bool Async = false>
void Sort_1Block_v6_register_batched(T* data, int n, int64_t arrays, hipStream_t stream = nullptr) {
    static_assert(Working / Threads <= 16, "v6 limits register tiles to 16 values per thread");
    if (int(Working) != working_length(n)) check(CS_E_UNSUPPORTED, "Sort_1Block_v6_register_batched (Working != W(n))");
    check(cs_batched_sort(DType<T>::value, data, data, arrays, n, CS_BITONIC_V6_REGISTER, int(Threads),
                          flags<Sorted, Shared, Async>(), stream),
          "Sort_1Block_v6_register_batched");
}

// Out-of-place form of any method (in and out may be the same pointer).
template <typename T>
void SortBatched(const T* in, T* out, int n, int64_t arrays, int method = CS_METHOD_AUTO, int threads = 0,
                 uint32_t flag_bits = 0, hipStream_t stream = nullptr) {
    check(cs_batched_sort(DType<T>::value, in, out, arrays, n, method, threads, flag_bits, stream), "SortBatched");
}

// Keys and int32 payloads in two arrays, `arrays` arrays of n each (no
// counterpart in the reference: the structure-of-arrays form of IntAndData
// and of the int64 + int32 record; same order and tie rule as sorting those
// records).  K is int or a 64-bit signed integer.
template <typename K>
void SortPairs_batched(K* keys, int* vals, int n, int64_t arrays, hipStream_t stream = nullptr) {
    static_assert(DType<K>::value == CS_I32 || DType<K>::value == CS_I64, "int32 or int64 keys");
    check(cs_batched_sort_pairs(DType<K>::value, keys, vals, keys, vals, arrays, n, stream), "SortPairs_batched");
}

}  // namespace cusort_hip

#endif
