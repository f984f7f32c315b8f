// range_alloc.h -- first-fit allocator of ranges of [0, cap): the device
// loop's arenas (bucket entries, window records, cursors, tile counters,
// pinned request records) are carved from buffers allocated once, so no
// allocation happens while the loop kernel runs.  Host only, no HIP.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <iterator>
#include <map>

namespace voss {
namespace serving {

struct RangeAlloc {
    uint64_t cap = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length, no two adjacent

    void reset(uint64_t c) {
        cap = c;
        free_.clear();
        if (c) free_[0] = c;
    }

    // The first free range that holds n units at an offset that is a
    // multiple of align; false if none does.  n = 0 always succeeds.
    bool alloc(uint64_t n, uint64_t align, uint64_t* off) {
        if (n == 0) {
            *off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            const uint64_t o = (it->first + align - 1) / align * align;
            const uint64_t end = it->first + it->second;
            if (o + n > end) continue;
            const uint64_t f0 = it->first;
            free_.erase(it);
            if (o > f0) free_[f0] = o - f0;
            if (o + n < end) free_[o + n] = end - (o + n);
            *off = o;
            return true;
        }
        return false;
    }

    // Returns [off, off + n) (allocated earlier) and merges it with its
    // free neighbours.
    void release(uint64_t off, uint64_t n) {
        if (n == 0) return;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

    uint64_t largest() const {
        uint64_t m = 0;
        for (const auto& f : free_) m = std::max(m, f.second);
        return m;
    }

    uint64_t free_total() const {
        uint64_t t = 0;
        for (const auto& f : free_) t += f.second;
        return t;
    }
};

}  // namespace serving
}  // namespace voss
