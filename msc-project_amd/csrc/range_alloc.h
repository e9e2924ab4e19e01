// range_alloc.h -- first-fit allocator of index ranges [0, capacity) with coalescing frees.
//
// The serving loop carves each admitted request's per-sample state and framebuffer out of
// fixed device arenas sized by the memory budget; this is the host-side bookkeeping of those
// arenas (no device calls), so admission never calls hipMalloc/hipFree, which would stall the
// running loop.
#pragma once

#include <cstdint>
#include <iterator>
#include <map>

namespace swr {

class RangeAlloc {
 public:
  explicit RangeAlloc(uint64_t capacity = 0) { reset(capacity); }

  void reset(uint64_t capacity) {
    free_.clear();
    capacity_ = capacity;
    used_ = 0;
    if (capacity) free_[0] = capacity;
  }

  // Lowest-addressed free range of n units; returns false when none is large enough.
  bool alloc(uint64_t n, uint64_t* offset) {
    if (n == 0) {
      *offset = 0;
      return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const uint64_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[off + n] = len - n;
      used_ += n;
      *offset = off;
      return true;
    }
    return false;
  }

  // Returns [offset, offset + n) and merges it with free neighbours.
  void release(uint64_t offset, uint64_t n) {
    if (n == 0) return;
    used_ -= n;
    auto next = free_.lower_bound(offset);
    uint64_t off = offset, len = n;
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && off + len == next->first) {
      len += next->second;
      free_.erase(next);
    }
    free_[off] = len;
  }

  bool fits(uint64_t n) const {
    for (const auto& kv : free_)
      if (kv.second >= n) return true;
    return n == 0;
  }
  uint64_t capacity() const { return capacity_; }
  uint64_t used() const { return used_; }
  size_t fragments() const { return free_.size(); }

 private:
  std::map<uint64_t, uint64_t> free_;  // offset -> length of each free range
  uint64_t capacity_ = 0, used_ = 0;
};

}  // namespace swr
