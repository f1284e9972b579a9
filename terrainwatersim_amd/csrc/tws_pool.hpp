// tws_pool.hpp -- first-fit sub-allocator over one device allocation.
//
// A batch reserves its HBM budget once (one hipMalloc) and carves grid
// states out of it, so adding and removing grids while the GPU runs never
// calls hipMalloc/hipFree, and admission is a plain "does a block of this
// size fit" question.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace tws {

class DevicePool {
public:
    static constexpr size_t kAlign = 256;

    // Takes ownership of [base, base + capacity).
    void reset(char* base, size_t capacity) {
        base_ = base;
        capacity_ = capacity;
        used_ = 0;
        free_.clear();
        if (capacity) free_[0] = capacity;
    }
    char* base() const { return base_; }
    size_t capacity() const { return capacity_; }
    size_t used() const { return used_; }

    static size_t round(size_t bytes) { return (bytes + kAlign - 1) / kAlign * kAlign; }

    // Returns the offset of a block of round(bytes), or SIZE_MAX if none fits.
    size_t alloc(size_t bytes) {
        const size_t need = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const size_t off = it->first, left = it->second - need;
            free_.erase(it);
            if (left) free_[off + need] = left;
            used_ += need;
            return off;
        }
        return SIZE_MAX;
    }

    // Returns a block to the pool, merging with free neighbours.
    void release(size_t off, size_t bytes) {
        size_t size = round(bytes);
        used_ -= size;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + size == next->first) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += size;
                return;
            }
        }
        free_[off] = size;
    }

    size_t largest_free() const {
        size_t m = 0;
        for (const auto& kv : free_) m = kv.second > m ? kv.second : m;
        return m;
    }

private:
    char* base_ = nullptr;
    size_t capacity_ = 0;
    size_t used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> size
};

}  // namespace tws
