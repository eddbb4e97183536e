// First-fit allocator of byte ranges inside one fixed device allocation
// (the serving runtime's HBM budget). Host-only bookkeeping: offsets, no
// device calls, so it is unit-tested on the CPU. Not thread-safe; the
// caller holds its lock.
#pragma once
#include <cstddef>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <vector>

namespace psxh {

class arena {
public:
    static constexpr size_t ALIGN = 256;
    static constexpr size_t NONE = ~size_t(0);

    explicit arena(size_t bytes = 0) { reset(bytes); }

    void reset(size_t bytes) {
        total_ = bytes & ~(ALIGN - 1);
        free_.clear();
        used_.clear();
        in_use_ = peak_ = 0;
        if (total_) free_[0] = total_;
    }

    static size_t round(size_t n) { return n == 0 ? ALIGN : (n + ALIGN - 1) & ~(ALIGN - 1); }

    // Offset of a block of at least n bytes, or NONE when no free range fits.
    size_t take(size_t n) {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_[off] = n;
            in_use_ += n;
            if (in_use_ > peak_) peak_ = in_use_;
            return off;
        }
        return NONE;
    }

    // Returns a block from take(); coalesces with its free neighbours.
    // False when off is not the start of a block in use.
    bool give(size_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        size_t len = u->second;
        used_.erase(u);
        in_use_ -= len;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + len) {
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

    size_t total() const { return total_; }
    size_t in_use() const { return in_use_; }
    size_t peak() const { return peak_; }
    size_t largest_free() const {
        size_t m = 0;
        for (const auto& f : free_) m = f.second > m ? f.second : m;
        return m;
    }
    size_t free_ranges() const { return free_.size(); }

private:
    size_t total_ = 0, in_use_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    std::map<size_t, size_t> used_;  // offset -> length
};

// Admission of queued requests into one launch: takes requests from the
// front of the queue, in arrival order, while their outputs fit in the arena
// and the batch holds fewer than max_batch; the first request that does not
// fit ends the batch (later, smaller requests do not overtake it). size_of(q)
// gives a request's bytes; set_off(q, off) records its block.
template <class Q, class SizeOf, class SetOff>
size_t admit_fifo(std::deque<Q>& pending, arena& ar, size_t max_batch, std::vector<Q>& batch, SizeOf size_of,
                  SetOff set_off) {
    size_t n = 0;
    while (!pending.empty() && batch.size() < max_batch) {
        Q q = pending.front();
        const size_t off = ar.take(size_of(q));
        if (off == arena::NONE) break;
        set_off(q, off);
        pending.pop_front();
        batch.push_back(q);
        ++n;
    }
    return n;
}

}  // namespace psxh
