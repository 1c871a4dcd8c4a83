// Ragged sort of segments of any length, for callers that manage their own
// device buffers (the serving layer): segments longer than the on-chip tile
// are cut into chunks, all chunks are sorted by one persistent launch, and
// the chunks of each long segment are merged by tiled merge passes.
#pragma once

#include <vector>

#include "partition.hpp"
#include "segmented.hpp"

namespace cs {

// Launch shape of the two persistent kernels over a bucket-ordered item
// list: items [0, split) belong to the multi-wave kernel, the rest to the
// single-wave one.
struct SegShape {
    int split = 0;
    int grid[2] = {0, 0};
};

struct RaggedPrep {
    std::vector<int64_t> chunk_off;   // CSR of the chunks (relative to the batch base)
    std::vector<int> list;            // chunk ids, bucket ordered
    std::vector<SegItem> items;       // persistent work items
    std::vector<MergeTile> tiles;     // merge tiles of the long segments
    int passes = 0;                   // merge passes needed by the longest segment
    SegShape shape;                   // persistent launches
};

// Builds the chunk CSR, work items and merge tiles for `segs` segments with
// CSR offsets h_off (host).  Returns CS_OK or an error code.
int ragged_prepare(int dtype, const int64_t* h_off, int64_t segs, RaggedPrep* p);

// Enqueues the sort: in -> out over all chunks, then merge passes between
// out and scratch (same layout as out; only touched when tiles exist).
// d_counters: two ints of device memory.
hipError_t ragged_launch(int dtype, const void* in, void* out, void* scratch, const int64_t* d_chunk_off,
                         const int* d_list, const SegItem* d_items, int nitems, int* d_counters,
                         const SegShape& shape, MergeTile* d_tiles, int ntiles, int passes, hipStream_t s);

// Bucketing shared with cs_segplan: list and items for CSR offsets whose
// segments all fit one tile; force_variant (per bucket) may be null.
int seg_build_items(int dtype, const int64_t* h_off, int64_t segs, const int* force_variant, std::vector<int>& list,
                    std::vector<SegItem>& items, int64_t* max_len, int* variants, int64_t* bucket_segments);
int seg_persistent_grid(int dtype, int which, int nitems);
// Per length class (bucket) b <= maxb: the variant seg_build_items picks and
// its segments per work item.
struct SegClasses {
    int maxb = 0;
    int small_max = 0;
    int variant[16] = {};
    int spi[16] = {};
};
SegClasses seg_classes(int dtype);
SegShape seg_shape(int dtype, const std::vector<SegItem>& items);
// Counter reset and both persistent launches, stream ordered.  With d_range
// (two ints in device memory: items of the multi-wave kernel, all items)
// the item ranges come from a plan built on the device and `sh` only gives
// the grid sizes.  reset = false: the caller has already zeroed d_counters
// in stream order.
hipError_t seg_launch_all(int dtype, const SegShape& sh, const void* in, void* out, const int64_t* off,
                          const int* d_list, const SegItem* d_items, int nitems, int* d_counters, hipStream_t s,
                          const int* d_range = nullptr, bool reset = true);

}  // namespace cs
