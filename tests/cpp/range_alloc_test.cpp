// Randomised check of the serving loop's arena allocator against a brute-force occupancy map.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "range_alloc.h"

int main() {
  std::mt19937_64 rng(12345);
  for (int round = 0; round < 50; ++round) {
    const uint64_t cap = 1 + rng() % 5000;
    swr::RangeAlloc a(cap);
    std::vector<int> owner(cap, -1);
    struct Live { uint64_t off, n; };
    std::vector<Live> live;
    for (int step = 0; step < 2000; ++step) {
      if (live.empty() || rng() % 3) {
        const uint64_t n = 1 + rng() % (cap / 4 + 1);
        uint64_t off = 0;
        // brute force: lowest free run of n
        long expect = -1;
        for (uint64_t i = 0, run = 0; i < cap; ++i) {
          run = owner[i] < 0 ? run + 1 : 0;
          if (run == n) { expect = (long)(i + 1 - n); break; }
        }
        const bool ok = a.alloc(n, &off);
        if (ok != (expect >= 0) || (ok && (long)off != expect)) {
          std::printf("FAIL round %d step %d: n=%llu ok=%d off=%llu expect=%ld\n", round, step,
                      (unsigned long long)n, ok, (unsigned long long)off, expect);
          return 1;
        }
        if (ok) {
          for (uint64_t i = off; i < off + n; ++i) owner[i] = (int)live.size();
          live.push_back({off, n});
        }
      } else {
        const size_t k = rng() % live.size();
        a.release(live[k].off, live[k].n);
        for (uint64_t i = live[k].off; i < live[k].off + live[k].n; ++i) owner[i] = -1;
        live.erase(live.begin() + k);
        for (uint64_t i = 0; i < cap; ++i)
          if (owner[i] >= 0) owner[i] = 0;  // ownership identity is not needed, only occupancy
      }
      uint64_t used = 0;
      for (int o : owner) used += o >= 0;
      if (used != a.used()) { std::printf("FAIL used %llu vs %llu\n", (unsigned long long)used, (unsigned long long)a.used()); return 1; }
    }
    for (auto& l : live) a.release(l.off, l.n);
    if (a.used() != 0 || a.fragments() != 1) { std::printf("FAIL: not coalesced (%zu)\n", a.fragments()); return 1; }
  }
  std::printf("ok\n");
  return 0;
}
