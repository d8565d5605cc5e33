// CPU check of the serving arena (era-boojum-cuda_amd/csrc/arena.hpp):
// random allocations (first fit and top end) and piecewise releases against a shadow list of live
// ranges; after every operation the free list must satisfy its invariants and
// no two live ranges may overlap.  Prints "OK <ops>" on success.
#include "../../era-boojum-cuda_amd/csrc/arena.hpp"

#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

struct Live {
  uint64_t off, len;
};

static int fail(const char *what, int step) {
  printf("FAIL %s at step %d\n", what, step);
  return 1;
}

int main() {
  const uint64_t align = 64 * 1024, cap = 1000 * align + 123;  // capacity rounds down
  bc::Arena a(cap, align);
  if (a.capacity() != 1000 * align || !a.check()) return fail("init", -1);
  uint64_t off = 0;
  // exact fill, then overflow
  if (!a.alloc(1000 * align, &off) || off != 0 || a.alloc(1, &off)) return fail("full", -1);
  a.release(0, 1000 * align);
  // alloc_high takes the top end: two pieces meet in the middle
  uint64_t hi = 0;
  if (!a.alloc_high(600 * align, &hi) || hi != 400 * align) return fail("high top end", -1);
  if (!a.alloc(400 * align, &off) || off != 0 || a.alloc_high(1, &off)) return fail("high fill", -1);
  a.release(0, 400 * align);
  a.release(hi, 600 * align);
  if (!a.check() || a.largest_free() != 1000 * align || a.in_use() != 0) return fail("empty", -1);

  std::mt19937_64 rng(7);
  std::vector<Live> live;
  const int steps = 200000;
  for (int s = 0; s < steps; ++s) {
    const int op = (int)(rng() % 3);
    if (op < 2 || live.empty()) {
      const uint64_t bytes = 1 + rng() % (40 * align);
      const bool high = rng() % 2;  // top end of the highest fitting range, or first fit
      if (high ? a.alloc_high(bytes, &off) : a.alloc(bytes, &off)) {
        const uint64_t len = a.round(bytes);
        if (off % align || off + len > a.capacity()) return fail("range", s);
        for (const Live &l : live)
          if (off < l.off + l.len && l.off < off + len) return fail("overlap", s);
        live.push_back({off, len});
      } else if (a.largest_free() >= a.round(bytes)) {
        return fail("spurious failure", s);
      }
    } else {
      const size_t k = rng() % live.size();
      Live l = live[k];
      if (l.len > align && rng() % 2) {  // release the tail piece only
        const uint64_t head = align * (1 + rng() % (l.len / align - 1));
        a.release(l.off + head, l.len - head);
        live[k].len = head;
      } else {
        a.release(l.off, l.len);
        live[k] = live.back();
        live.pop_back();
      }
    }
    uint64_t used = 0;
    for (const Live &l : live) used += l.len;
    if (!a.check() || used != a.in_use()) return fail("invariants", s);
  }
  for (const Live &l : live) a.release(l.off, l.len);
  if (!a.check() || a.largest_free() != a.capacity() || a.in_use() != 0) return fail("drain", steps);
  printf("OK %d high_water=%llu\n", steps, (unsigned long long)a.high_water());
  return 0;
}
