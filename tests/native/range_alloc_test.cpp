// Unit test of the request loop's pool allocator (host only).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../hiqsimulator_amd/csrc/range_alloc.hpp"

#define CHECK(c)                                                         \
     do {                                                                \
          if (!(c)) {                                                    \
               std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
               return 1;                                                 \
          }                                                              \
     } while (0)

int main()
{
     using hq::RangeAlloc;
     RangeAlloc a(1000);
     uint64_t o1, o2, o3, o4;
     CHECK(a.alloc(100, 1, &o1) && o1 == 0);
     CHECK(a.alloc(100, 64, &o2) && o2 == 128);  // aligned: the gap 100..128 stays free
     CHECK(a.free_total() == 800);
     CHECK(a.alloc(28, 1, &o3) && o3 == 100);     // first fit takes the gap
     CHECK(!a.alloc(800, 1, &o4));                // 772 left
     a.release(o2, 100);
     a.release(o1, 100);
     a.release(o3, 28);
     CHECK(a.free_total() == 1000 && a.free_ranges() == 1);  // fully coalesced
     CHECK(a.alloc(1000, 8, &o4) && o4 == 0);
     CHECK(!a.alloc(1, 1, &o4));
     a.release(0, 1000);
     // random alloc/free against a byte map
     std::mt19937 rng(5);
     const uint64_t cap = 1 << 16;
     RangeAlloc r(cap);
     std::vector<int> owner(cap, -1);
     std::vector<std::pair<uint64_t, uint64_t>> live;
     for (int step = 0; step < 20000; ++step) {
          if (live.empty() || rng() % 3) {
               const uint64_t size = 1 + rng() % 700, align = uint64_t(1) << (rng() % 7);
               uint64_t off;
               if (r.alloc(size, align, &off)) {
                    CHECK(off % align == 0 && off + size <= cap);
                    for (uint64_t i = off; i < off + size; ++i) {
                         CHECK(owner[i] < 0);
                         owner[i] = step;
                    }
                    live.push_back({off, size});
               }
          }
          else {
               const size_t k = rng() % live.size();
               for (uint64_t i = live[k].first; i < live[k].first + live[k].second; ++i) owner[i] = -1;
               r.release(live[k].first, live[k].second);
               live[k] = live.back();
               live.pop_back();
          }
          uint64_t used = 0;
          for (auto& l: live) used += l.second;
          CHECK(r.free_total() == cap - used);
     }
     for (auto& l: live) r.release(l.first, l.second);
     CHECK(r.free_total() == cap && r.free_ranges() == 1);
     std::printf("range_alloc ok\n");
     return 0;
}
