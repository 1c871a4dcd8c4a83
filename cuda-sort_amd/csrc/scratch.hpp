// Device temporaries of one library call.  With an arena (a region of a
// cs_workspace, sized when the workspace is reserved) every temporary is
// carved from it in order and nothing is allocated; a request beyond the
// arena fails with CS_E_WORKSPACE rather than falling back to an allocation.
// Without an arena each temporary is a stream-ordered allocation from the
// library's pool, released in stream order when the Scratch goes out of
// scope.  Arena users must not run two calls on one arena concurrently (the
// workspace orders its users through an event).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/cusort_hip.h"
#include "gen.hpp"

namespace cs {

constexpr size_t kScratchAlign = 256;
constexpr size_t scratch_round(size_t bytes) { return (bytes + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }

class Scratch {
public:
    explicit Scratch(hipStream_t s, void* arena = nullptr, size_t arena_bytes = 0)
        : s_(s), base_(static_cast<char*>(arena)), cap_(arena ? arena_bytes : 0) {}
    ~Scratch() {
        for (void* p : owned_) (void)hipFreeAsync(p, s_);
    }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;

    // CS_OK, CS_E_WORKSPACE (arena exhausted) or a hipError_t.
    template <class T>
    int get(T** p, size_t count) {
        const size_t bytes = scratch_round((count ? count : 1) * sizeof(T));
        if (base_) {
            if (used_ + bytes > cap_) return CS_E_WORKSPACE;
            *p = reinterpret_cast<T*>(base_ + used_);
            used_ += bytes;
            return CS_OK;
        }
        void* q = nullptr;
        const hipError_t e = pool_malloc_async(&q, bytes, s_);
        if (e != hipSuccess) return int(e);
        owned_.push_back(q);
        *p = static_cast<T*>(q);
        used_ += bytes;
        return CS_OK;
    }
    bool from_arena() const { return base_ != nullptr; }
    size_t used() const { return used_; }
    hipStream_t stream() const { return s_; }
    // Upper bound of the persistent grids the call launches (a workspace's
    // blocks per CU times the CU count; unbounded otherwise).
    int grid_cap = 1 << 30;

private:
    hipStream_t s_;
    char* base_;
    size_t cap_;
    size_t used_ = 0;
    std::vector<void*> owned_;
};

// Upper bounds of the temporaries of one sort of n elements (bytes, each
// temporary rounded up to kScratchAlign), used to size workspace arenas.
size_t merge_scratch_bytes(int dtype, int64_t n);
size_t partition_scratch_bytes(int dtype, int64_t n, int min_parts, int max_parts);

// Chunk sort + merge passes (large.cpp): result in `final` (== data for in
// place); `scratch_elems` holds at least n elements and may be `data` itself
// for an out-of-place call.
int merge_sort(int dtype, void* data, int64_t n, void* final, void* scratch_elems, Scratch& tmp);

// Sample-sort partition (partition.hip): sorts data[0, n) into out with
// `values` (>= n elements) as scratch.
int partition_sort(int dtype, void* data, int64_t n, void* out, void* values, size_t values_elems, int min_parts,
                   int max_parts, Scratch& tmp);

}  // namespace cs
