// hipCUB / rocThrust baselines for cusort_bench (see baselines.hpp).
#include "baselines.hpp"

#include <climits>

#include <hipcub/hipcub.hpp>
#include <thrust/device_ptr.h>
#include <thrust/execution_policy.h>
#include <thrust/sort.h>
#include <thrust/version.h>

namespace csbench {
namespace {

struct Rec {
    int32_t key;
    int32_t payload;
};

// Padding must stay strictly after valid records even for INT_MAX keys: the
// generated payloads are < INT_MAX, the padding record's is INT_MAX (the
// reference's CubLess, bench.cu:396-404).
struct RecLess {
    __host__ __device__ bool operator()(const Rec& a, const Rec& b) const {
        if (a.key == b.key) return a.payload != INT_MAX && b.payload == INT_MAX;
        return a.key < b.key;
    }
};
struct IntLess {
    __host__ __device__ bool operator()(int32_t a, int32_t b) const { return a < b; }
};

template <class T> struct Traits;
template <> struct Traits<int32_t> {
    using Less = IntLess;
    __host__ __device__ static int32_t pad() { return INT_MAX; }
};
template <> struct Traits<Rec> {
    using Less = RecLess;
    __host__ __device__ static Rec pad() { return Rec{INT_MAX, INT_MAX}; }
};

template <class T, int Threads, int Items>
__global__ void __launch_bounds__(Threads) k_block_merge(T* data, int n) {
    using Load = hipcub::BlockLoad<T, Threads, Items, hipcub::BLOCK_LOAD_WARP_TRANSPOSE>;
    using Store = hipcub::BlockStore<T, Threads, Items, hipcub::BLOCK_STORE_WARP_TRANSPOSE>;
    using Sort = hipcub::BlockMergeSort<T, Threads, Items>;
    __shared__ union {
        typename Load::TempStorage load;
        typename Store::TempStorage store;
        typename Sort::TempStorage sort;
    } temp;
    T keys[Items];
    T* segment = data + size_t(blockIdx.x) * n;
    Load(temp.load).Load(segment, keys, n, Traits<T>::pad());
    __syncthreads();
    Sort(temp.sort).Sort(keys, typename Traits<T>::Less{}, n, Traits<T>::pad());
    __syncthreads();
    Store(temp.store).Store(segment, keys, n);
}

using Launcher = bool (*)(void*, int, int, hipStream_t, KernelAttrs*);

template <class T, int Threads, int Items>
bool launch(void* data, int n, int arrays, hipStream_t s, KernelAttrs* attrs) {
    const void* fn = reinterpret_cast<const void*>(&k_block_merge<T, Threads, Items>);
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, fn) != hipSuccess) { (void)hipGetLastError(); return false; }
    int active = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&active, fn, Threads, 0) != hipSuccess || active <= 0) {
        (void)hipGetLastError();
        return false;
    }
    if (attrs) {
        attrs->registers = fa.numRegs;
        attrs->active_blocks_per_cu = active;
        attrs->local_bytes = static_cast<long long>(fa.localSizeBytes);
        attrs->shared_bytes = fa.sharedSizeBytes;
    }
    if (arrays > 0)
        hipLaunchKernelGGL((k_block_merge<T, Threads, Items>), dim3(unsigned(arrays)), dim3(Threads), 0, s,
                           static_cast<T*>(data), n);
    return hipGetLastError() == hipSuccess;
}

struct Entry {
    int type, threads, items;
    Launcher fn;
};

// Threads 64..1024, items 1..32, up to the 16384-element (int) / 8192 (record) capacity.
#define CS_BM(T, TY, TH, IT) Entry{TY, TH, IT, &launch<T, TH, IT>}
#define CS_BM_ROW(T, TY, TH) CS_BM(T, TY, TH, 1), CS_BM(T, TY, TH, 2), CS_BM(T, TY, TH, 4), CS_BM(T, TY, TH, 8), \
                             CS_BM(T, TY, TH, 16)
const Entry kEntries[] = {
    CS_BM_ROW(int32_t, 0, 64),  CS_BM(int32_t, 0, 64, 32),  CS_BM_ROW(int32_t, 0, 128), CS_BM(int32_t, 0, 128, 32),
    CS_BM_ROW(int32_t, 0, 256), CS_BM(int32_t, 0, 256, 32), CS_BM_ROW(int32_t, 0, 512), CS_BM(int32_t, 0, 512, 32),
    CS_BM_ROW(int32_t, 0, 1024),
    CS_BM_ROW(Rec, 1, 64),  CS_BM(Rec, 1, 64, 32),  CS_BM_ROW(Rec, 1, 128), CS_BM(Rec, 1, 128, 32),
    CS_BM_ROW(Rec, 1, 256), CS_BM(Rec, 1, 256, 32), CS_BM_ROW(Rec, 1, 512),
    CS_BM(Rec, 1, 1024, 1), CS_BM(Rec, 1, 1024, 2), CS_BM(Rec, 1, 1024, 4), CS_BM(Rec, 1, 1024, 8),
};
#undef CS_BM_ROW
#undef CS_BM

}  // namespace

std::vector<BlockMergeConfig> block_merge_configs(int type) {
    std::vector<BlockMergeConfig> v;
    for (const Entry& e : kEntries)
        if (e.type == type) v.push_back({e.threads, e.items});
    return v;
}

bool block_merge_sort(int type, int threads, int items, void* data, int n, int arrays, hipStream_t stream,
                      KernelAttrs* attrs) {
    for (const Entry& e : kEntries)
        if (e.type == type && e.threads == threads && e.items == items && threads * items >= n)
            return e.fn(data, n, arrays, stream, attrs);
    return false;
}

void thrust_segment_loop(int type, void* data, int n, int arrays, hipStream_t stream) {
    for (int a = 0; a < arrays; ++a) {
        if (type == 0) {
            auto first = thrust::device_pointer_cast(static_cast<int32_t*>(data) + size_t(a) * n);
            thrust::sort(thrust::hip::par.on(stream), first, first + n);
        } else {
            auto first = thrust::device_pointer_cast(static_cast<Rec*>(data) + size_t(a) * n);
            thrust::sort(thrust::hip::par.on(stream), first, first + n, RecLess{});
        }
    }
}

int hipcub_version() { return HIPCUB_VERSION; }
int thrust_version() { return THRUST_VERSION; }

}  // namespace csbench
