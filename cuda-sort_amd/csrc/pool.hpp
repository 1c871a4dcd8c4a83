// Offset allocator for the serving layer's device pool: first fit over an
// address-ordered free list, neighbours coalesced on release, every block
// aligned.  Plain C++ with no HIP dependency (the tests exercise it on the
// host); the server maps offsets onto one device allocation whose size is
// the HBM admission budget.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace cs {

class OffsetPool {
public:
    static constexpr size_t kAlign = 256;

    explicit OffsetPool(size_t capacity = 0) { reset(capacity); }

    void reset(size_t capacity) {
        free_.clear();
        used_.clear();
        capacity_ = capacity / kAlign * kAlign;
        in_use_ = 0;
        peak_ = 0;
        if (capacity_) free_[0] = capacity_;
    }

    static size_t round_up(size_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }

    // Returns true and the offset of a block of at least `bytes`.
    bool alloc(size_t bytes, size_t* offset) {
        const size_t need = round_up(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= need) {
                const size_t off = it->first;
                const size_t rest = it->second - need;
                free_.erase(it);
                if (rest) free_[off + need] = rest;
                used_[off] = need;
                in_use_ += need;
                if (in_use_ > peak_) peak_ = in_use_;
                *offset = off;
                return true;
            }
        }
        return false;
    }

    // Releases a block returned by alloc; false for an unknown offset.
    bool release(size_t offset) {
        auto u = used_.find(offset);
        if (u == used_.end()) return false;
        size_t off = offset, size = u->second;
        used_.erase(u);
        in_use_ -= size;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + size == next->first) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += size;
                return true;
            }
        }
        free_[off] = size;
        return true;
    }

    bool fits_ever(size_t bytes) const { return round_up(bytes ? bytes : 1) <= capacity_; }
    size_t capacity() const { return capacity_; }
    size_t in_use() const { return in_use_; }
    size_t peak() const { return peak_; }
    size_t largest_free() const {
        size_t m = 0;
        for (const auto& f : free_) m = f.second > m ? f.second : m;
        return m;
    }
    size_t free_blocks() const { return free_.size(); }
    size_t used_blocks() const { return used_.size(); }

private:
    std::map<size_t, size_t> free_;   // offset -> size, address ordered
    std::map<size_t, size_t> used_;   // offset -> size
    size_t capacity_ = 0;
    size_t in_use_ = 0;
    size_t peak_ = 0;
};

}  // namespace cs
