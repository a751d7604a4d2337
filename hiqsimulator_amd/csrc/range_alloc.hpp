// First-fit range allocator over [0, capacity) with coalescing of freed
// ranges (host only; the request loop carves its device pools with it).
#pragma once

#include <cstdint>
#include <iterator>
#include <map>

namespace hq {

class RangeAlloc {
public:
     explicit RangeAlloc(uint64_t capacity = 0)
     {
          if (capacity) free_[0] = capacity;
          free_total_ = capacity;
     }

     // First free range that holds `size` units at an `align` multiple.
     bool alloc(uint64_t size, uint64_t align, uint64_t* off)
     {
          if (size == 0) size = 1;
          if (align == 0) align = 1;
          for (auto it = free_.begin(); it != free_.end(); ++it) {
               const uint64_t start = (it->first + align - 1) / align * align;
               const uint64_t end = it->first + it->second;
               if (start + size > end) continue;
               const uint64_t f_off = it->first, f_end = end;
               free_.erase(it);
               if (start > f_off) free_[f_off] = start - f_off;
               if (start + size < f_end) free_[start + size] = f_end - (start + size);
               free_total_ -= size;
               *off = start;
               return true;
          }
          return false;
     }

     // Returns [off, off + size) and merges it with free neighbours.
     void release(uint64_t off, uint64_t size)
     {
          if (size == 0) size = 1;
          auto next = free_.lower_bound(off);
          uint64_t start = off, end = off + size;
          if (next != free_.begin()) {
               auto prev = std::prev(next);
               if (prev->first + prev->second == start) {
                    start = prev->first;
                    free_.erase(prev);
               }
          }
          if (next != free_.end() && next->first == end) {
               end = next->first + next->second;
               free_.erase(next);
          }
          free_[start] = end - start;
          free_total_ += size;
     }

     uint64_t free_total() const { return free_total_; }
     size_t free_ranges() const { return free_.size(); }

private:
     std::map<uint64_t, uint64_t> free_;  // offset -> size
     uint64_t free_total_ = 0;
};

}  // namespace hq
