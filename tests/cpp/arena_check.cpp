// Randomised check of the serving arena's sub-allocator (pme_arena.hpp):
// allocations, releases and the evict-until-the-freed-block-fits sequence of
// the serving loop's admission, with the allocator's invariants checked after
// every step against a list of the live blocks.  Prints "ok" and exits 0, or
// names the first violated invariant and exits 1.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "pme_arena.hpp"

using pme::Arena;

struct Block {
  size_t off, bytes;
};

static size_t granule(size_t b) { return (b + 255) & ~(size_t)255; }

static bool fail(const char* what, int step) {
  std::printf("step %d: %s\n", step, what);
  return false;
}

static bool check(const Arena& a, const std::vector<Block>& live, int step) {
  size_t used = 0;
  for (size_t i = 0; i < live.size(); i++) {
    const Block& x = live[i];
    if (x.off % 256 || x.off + granule(x.bytes) > a.size) return fail("block outside the arena or misaligned", step);
    used += granule(x.bytes);
    for (size_t j = i + 1; j < live.size(); j++) {
      const Block& y = live[j];
      if (x.off < y.off + granule(y.bytes) && y.off < x.off + granule(x.bytes)) return fail("live blocks overlap", step);
    }
    for (const auto& f : a.free_)
      if (x.off < f.first + f.second && f.first < x.off + granule(x.bytes)) return fail("free range overlaps a live block", step);
  }
  if (used != a.used) return fail("used bytes differ from the live blocks", step);
  size_t fsum = 0, prev_end = 0;
  bool first = true;
  for (const auto& f : a.free_) {
    if (f.second == 0) return fail("empty free range", step);
    if (!first && f.first <= prev_end) return fail("free ranges touch or overlap (not coalesced)", step);
    prev_end = f.first + f.second;
    first = false;
    fsum += f.second;
  }
  if (fsum + used != a.size) return fail("free + used != size", step);
  return true;
}

int main() {
  std::mt19937_64 rng(20261202);
  for (int trial = 0; trial < 20; trial++) {
    Arena a;
    a.init(((size_t)1 << 20) + 77);  // not a granule multiple: rounds down
    std::vector<Block> live;
    std::uniform_int_distribution<size_t> sz(1, 96 * 1024);
    for (int step = 0; step < 4000; step++) {
      const int op = (int)(rng() % 3);
      if (op < 2) {  // admit: first fit, else evict the oldest blocks until the freed one fits
        const size_t need = sz(rng);
        size_t off = 0;
        if (!a.alloc(need, &off)) {
          bool placed = false;
          while (!live.empty()) {
            const Block v = live.front();
            live.erase(live.begin());
            const std::pair<size_t, size_t> blk = a.release(v.off, v.bytes);
            if (a.free_.count(blk.first) == 0 || a.free_.at(blk.first) != blk.second)
              return fail("release did not return a free range", step), 1;
            if (!(blk.first <= v.off && v.off + granule(v.bytes) <= blk.first + blk.second))
              return fail("the returned range does not hold the released block", step), 1;
            if (blk.second >= granule(need)) {
              a.take(blk.first, need, &off);
              placed = true;
              break;
            }
            if (!check(a, live, step)) return 1;
          }
          if (!placed) {
            if (granule(need) <= a.size) return fail("an empty arena did not take the block", step), 1;
            continue;
          }
        }
        live.push_back(Block{off, need});
      } else if (!live.empty()) {  // release a random live block
        const size_t k = rng() % live.size();
        a.release(live[k].off, live[k].bytes);
        live.erase(live.begin() + (long)k);
      }
      if (!check(a, live, step)) return 1;
    }
  }
  std::printf("ok\n");
  return 0;
}
