// Batch admission of the serving layer: which pending requests go into the
// next batch.  Plain C++ with no HIP dependency, so the policy is tested on
// the host.
//
// The pending requests are scanned in arrival order (at most `window` of
// them) and each one is taken if the batch stays within the element and
// request caps and its device footprint within `room`, the largest free
// block of the pool.  The footprint is elems * esz, doubled when any taken
// request is longer than one chunk (the merge passes need a scratch copy).
// A request that does not fit is passed over, so later requests that fit go
// ahead of it, unless it has already been passed over `bypass_limit` times:
// then the scan stops there and nothing behind it is admitted until it fits,
// which bounds how long any request can wait behind later ones.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "pool.hpp"

namespace cs {

struct AdmitLimits {
    int64_t max_elems = 0;    // elements per batch
    int max_reqs = 0;         // requests per batch
    int64_t chunk = 0;        // longest request sorted without merge passes
    size_t esz = 0;           // bytes per element
    size_t room = 0;          // largest free block of the pool
    size_t window = 1024;     // pending requests examined
    int bypass_limit = 16;    // times a request may be passed over
};

struct AdmitResult {
    std::vector<size_t> taken;    // indices into the pending sequence, ascending
    std::vector<size_t> bypassed; // indices passed over by a later taken request
    int64_t elems = 0;
    bool long_req = false;
    size_t bytes = 0;             // pool bytes the batch needs
};

inline size_t admit_bytes(int64_t elems, size_t esz, bool long_req) {
    return size_t(elems) * esz * (long_req ? 2 : 1);
}

// size_of(i) and passes_of(i) give pending request i's length and how many
// times it has been passed over.
template <class SizeOf, class PassesOf>
AdmitResult select_batch(size_t pending, SizeOf size_of, PassesOf passes_of, const AdmitLimits& lim) {
    AdmitResult r;
    std::vector<size_t> passed;
    for (size_t i = 0; i < pending && i < lim.window; ++i) {
        if (int(r.taken.size()) >= lim.max_reqs) break;
        const int64_t n = size_of(i);
        const bool lr = r.long_req || n > lim.chunk;
        const size_t need = OffsetPool::round_up(admit_bytes(r.elems + n, lim.esz, lr));
        if (r.elems + n <= lim.max_elems && need <= lim.room) {
            r.elems += n;
            r.long_req = lr;
            r.taken.push_back(i);
            continue;
        }
        if (passes_of(i) >= lim.bypass_limit) break;
        passed.push_back(i);
    }
    if (!r.taken.empty()) {
        for (size_t i : passed)
            if (i < r.taken.back()) r.bypassed.push_back(i);
        r.bytes = admit_bytes(r.elems, lim.esz, r.long_req);
    }
    return r;
}

}  // namespace cs
