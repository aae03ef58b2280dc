// First-fit allocator of byte ranges inside one block of memory (offsets
// only; the caller owns the memory and chooses the alignment by rounding the
// sizes it passes). Adjacent free ranges are merged on release.
#pragma once

#include <cstdint>
#include <iterator>
#include <map>

namespace svo {

class RangeAlloc {
public:
    void reset(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(uint64_t bytes, uint64_t &off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }

private:
    std::map<uint64_t, uint64_t> free_;  // offset -> length
};

}  // namespace svo
