// arena.hpp — first-fit allocator over the byte offsets of one device
// allocation (the serving HBM budget). Host-only code with no HIP
// dependency, so its bookkeeping is unit-tested on the CPU.
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <map>

namespace bnmf {

class Arena {
   public:
    static constexpr size_t kAlign = 256;
    static size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

    explicit Arena(size_t cap) : cap_(align_up(cap)) {
        if (cap_) free_[0] = cap_;
    }
    // Lowest-offset free block that holds n bytes (rounded up to 256).
    bool alloc(size_t n, size_t* off) {
        n = align_up(std::max<size_t>(n, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n;
                const size_t noff = it->first + n;
                free_.erase(it);
                if (rest) free_[noff] = rest;
                used_ += n;
                return true;
            }
        }
        return false;
    }
    // Returns a block from alloc(); neighbouring free blocks merge.
    void release(size_t off, size_t n) {
        n = align_up(std::max<size_t>(n, 1));
        used_ -= n;
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
    bool fits(size_t n) const {
        n = align_up(std::max<size_t>(n, 1));
        for (const auto& kv : free_)
            if (kv.second >= n) return true;
        return false;
    }
    size_t used() const { return used_; }
    size_t capacity() const { return cap_; }
    size_t free_blocks() const { return free_.size(); }

   private:
    size_t cap_, used_ = 0;
    std::map<size_t, size_t> free_;
};

}  // namespace bnmf
