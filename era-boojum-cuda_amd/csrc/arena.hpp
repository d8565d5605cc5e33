// First-fit allocator over the byte range [0, capacity) of one HBM
// allocation (host bookkeeping only; no device calls).  Free ranges are kept
// in an ordered map and coalesced with both neighbours on release, so a
// range may be returned in pieces (e.g. a request's inputs at completion and
// its outputs later).  All sizes and offsets are multiples of `align`.
#pragma once
#include <stdint.h>

#include <iterator>
#include <map>

namespace bc {

class Arena {
 public:
  Arena(uint64_t capacity, uint64_t align) : align_(align), cap_(capacity / align * align) {
    if (cap_) free_[0] = cap_;
  }

  uint64_t round(uint64_t bytes) const { return (bytes + align_ - 1) / align_ * align_; }
  uint64_t capacity() const { return cap_; }
  uint64_t in_use() const { return used_; }
  uint64_t high_water() const { return high_; }

  // Lowest-addressed free range that holds `bytes`; false if none does.
  bool alloc(uint64_t bytes, uint64_t *off) {
    const uint64_t need = round(bytes ? bytes : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      *off = it->first;
      const uint64_t rest = it->second - need;
      const uint64_t rest_off = it->first + need;
      free_.erase(it);
      if (rest) free_[rest_off] = rest;
      used_ += need;
      if (used_ > high_) high_ = used_;
      return true;
    }
    return false;
  }

  // The top end of the highest-addressed free range that holds `bytes`;
  // false if none does.  Long-lived buffers placed this way stay clear of
  // the low end, where short-lived ones come and go.
  bool alloc_high(uint64_t bytes, uint64_t *off) {
    const uint64_t need = round(bytes ? bytes : 1);
    for (auto it = free_.rbegin(); it != free_.rend(); ++it) {
      if (it->second < need) continue;
      const uint64_t rest = it->second - need;
      *off = it->first + rest;
      if (rest) it->second = rest;
      else free_.erase(std::next(it).base());
      used_ += need;
      if (used_ > high_) high_ = used_;
      return true;
    }
    return false;
  }

  // Returns [off, off + round(bytes)) to the free list.  The range must be
  // allocated (a whole allocation or an aligned piece of one).
  void release(uint64_t off, uint64_t bytes) {
    uint64_t len = round(bytes ? bytes : 1);
    used_ -= len;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += len;
        return;
      }
    }
    free_[off] = len;
  }

  // Largest single free range.
  uint64_t largest_free() const {
    uint64_t m = 0;
    for (const auto &kv : free_)
      if (kv.second > m) m = kv.second;
    return m;
  }

  // Invariants: free ranges sorted, disjoint, never adjacent, inside the
  // capacity, and free + used == capacity.
  bool check() const {
    uint64_t end = 0, total = 0;
    bool first = true;
    for (const auto &kv : free_) {
      if (kv.second == 0 || kv.first % align_ || kv.second % align_) return false;
      if (!first && kv.first <= end) return false;  // overlapping or adjacent (not coalesced)
      end = kv.first + kv.second;
      if (end > cap_) return false;
      total += kv.second;
      first = false;
    }
    return total + used_ == cap_;
  }

 private:
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  uint64_t align_;
  uint64_t cap_;
  uint64_t used_ = 0;
  uint64_t high_ = 0;
};

}  // namespace bc
