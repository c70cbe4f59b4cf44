// ccfv_sched.h — host bookkeeping of the serving engine: a first-fit
// allocator over the engine's device arena and the admission queue.
//
// Nothing here touches the device (offsets, not pointers), so the policy is
// unit-tested on the CPU by a small C++ driver.
#pragma once

#include <cstddef>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <vector>

namespace ccfv {
namespace sched {

constexpr size_t NPOS = ~(size_t)0;

inline size_t round_up(size_t b, size_t a) { return (b + a - 1) / a * a; }

// Device bytes one resident Parareal request holds: U (ns+1 fields), G and F
// (ns each), the coarse work field, the per-block increment partials of the
// ns+1 slots (nb blocks per slot), each rounded to the arena alignment.
inline size_t request_bytes(int nx, int ny, int ns, int nb, size_t align = 256)
{
    const size_t f = sizeof(double) * (size_t)nx * ny;
    size_t b = round_up(f * (ns + 1), align);
    b += 2 * round_up(f * ns, align);
    b += round_up(f, align);
    b += round_up(sizeof(double) * 2 * (size_t)nb * (ns + 1), align);
    return b;
}

// First-fit allocator over [0, size): free blocks keyed by offset, merged with
// their neighbours on release.
class ArenaAlloc {
public:
    explicit ArenaAlloc(size_t size = 0, size_t align = 256) { reset(size, align); }

    void reset(size_t size, size_t align = 256)
    {
        align_ = align ? align : 1;
        size_ = size / align_ * align_;
        free_.clear();
        if (size_) free_[0] = size_;
        used_ = 0;
    }

    // offset of a block of at least `bytes`, or NPOS
    size_t alloc(size_t bytes)
    {
        bytes = round_up(bytes ? bytes : 1, align_);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, rem = it->second - bytes;
            free_.erase(it);
            if (rem) free_[off + bytes] = rem;
            used_ += bytes;
            return off;
        }
        return NPOS;
    }

    void release(size_t off, size_t bytes)
    {
        bytes = round_up(bytes ? bytes : 1, align_);
        auto it = free_.emplace(off, bytes).first;
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
        used_ -= bytes;
    }

    bool fits(size_t bytes) const { return largest_free() >= round_up(bytes ? bytes : 1, align_); }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t free_blocks() const { return free_.size(); }
    size_t largest_free() const
    {
        size_t m = 0;
        for (auto &kv : free_) m = kv.second > m ? kv.second : m;
        return m;
    }

private:
    size_t size_ = 0, align_ = 256, used_ = 0;
    std::map<size_t, size_t> free_;
};

// Waiting requests in arrival order.  admit() takes every request that fits
// (arena block and an engine slot), oldest first; a request that does not
// fit is passed over, but once the oldest waiting one has been passed over
// max_bypass times nothing behind it is admitted until it is, so a stream of
// small requests cannot starve a large one.
class AdmissionQueue {
public:
    struct Entry {
        int64_t id;
        size_t bytes;
        int bypassed;
    };
    struct Grant {
        int64_t id;
        size_t offset;
    };

    explicit AdmissionQueue(int max_bypass = 8) : max_bypass_(max_bypass) {}

    void push(int64_t id, size_t bytes) { q_.push_back(Entry{id, bytes, 0}); }
    size_t size() const { return q_.size(); }
    bool empty() const { return q_.empty(); }

    std::vector<Grant> admit(ArenaAlloc &arena, int free_slots)
    {
        std::vector<Grant> out;
        auto it = q_.begin();
        bool head_blocked = false;
        while (it != q_.end() && free_slots > 0) {
            const bool is_head = it == q_.begin();
            if (!is_head && head_blocked && q_.front().bypassed >= max_bypass_) break;
            const size_t off = arena.alloc(it->bytes);
            if (off == NPOS) {
                if (is_head) head_blocked = true;
                ++it;
                continue;
            }
            out.push_back(Grant{it->id, off});
            if (!is_head && head_blocked) q_.front().bypassed++;
            it = q_.erase(it);
            free_slots--;
            if (is_head) head_blocked = false;
        }
        return out;
    }

private:
    int max_bypass_;
    std::deque<Entry> q_;
};

}  // namespace sched
}  // namespace ccfv
