// sun_arena.hpp -- first-fit sub-allocator over one pre-allocated device region.
//
// The job server reserves its whole HBM budget once and carves every job's link
// field and scratch out of it, so admitting and retiring jobs never calls
// hipMalloc / hipFree (hipFree synchronises the device and would stall the jobs
// still running).  Host-only bookkeeping: offsets, no device access.  Freed blocks
// merge with free neighbours.  Not thread-safe; the server serialises access.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace sun {

class Arena {
   public:
    static constexpr size_t kAlign = 256;

    void reset(size_t capacity) {
        cap_ = capacity - capacity % kAlign;
        free_.clear();
        used_.clear();
        if (cap_) free_[0] = cap_;
        in_use_ = 0;
    }
    static size_t round_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

    // offset of a block of n bytes, or SIZE_MAX when no free block is large enough
    size_t alloc(size_t n) {
        if (n == 0) n = kAlign;
        n = round_up(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_[off] = n;
            in_use_ += n;
            return off;
        }
        return SIZE_MAX;
    }

    // returns false for an offset that is not an allocated block
    bool release(size_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        size_t len = u->second;
        used_.erase(u);
        in_use_ -= len;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return true;
            }
        }
        free_[off] = len;
        return true;
    }

    size_t capacity() const { return cap_; }
    size_t in_use() const { return in_use_; }
    size_t largest_free() const {
        size_t m = 0;
        for (auto& kv : free_) m = kv.second > m ? kv.second : m;
        return m;
    }
    size_t blocks() const { return used_.size(); }

   private:
    size_t cap_ = 0, in_use_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    std::map<size_t, size_t> used_;  // offset -> length
};

}  // namespace sun
