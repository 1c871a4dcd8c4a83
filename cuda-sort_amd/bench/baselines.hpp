// Vendor baselines of the benchmark driver, standing in for the reference's
// CUB BlockMergeSort and Thrust segment loop (bench.cu:395-423, 554-585):
// hipCUB BlockMergeSort (one workgroup per array) and rocThrust thrust::sort
// called once per array.  Element types: 0 = int32, 1 = IntAndData
// {int32 key, int32 payload}.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace csbench {

struct BlockMergeConfig {
    int threads;
    int items;
};

struct KernelAttrs {
    int registers = -1;
    int active_blocks_per_cu = -1;
    long long local_bytes = -1;
    size_t shared_bytes = 0;
};

// The (threads, items) pairs compiled for `type`, threads * items a power of two.
std::vector<BlockMergeConfig> block_merge_configs(int type);

// Sorts `arrays` arrays of n elements in place with hipCUB BlockMergeSort
// <threads, items> (threads * items >= n).  Returns false if that
// configuration was not compiled or cannot launch; attrs gets its resources.
bool block_merge_sort(int type, int threads, int items, void* data, int n, int arrays, hipStream_t stream,
                      KernelAttrs* attrs);

// thrust::sort on each array in turn (one library call per array).
void thrust_segment_loop(int type, void* data, int n, int arrays, hipStream_t stream);

int hipcub_version();
int thrust_version();

}  // namespace csbench
