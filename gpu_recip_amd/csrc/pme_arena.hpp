// pme_arena.hpp -- sub-allocator of the serving loop's HBM arena (host
// bookkeeping only: offsets into one device allocation made at creation, so
// admitting and evicting systems never calls the HIP allocator).
#pragma once
#include <cstddef>
#include <iterator>
#include <map>
#include <utility>

namespace pme {

// First-fit allocator over [0, size) with coalescing; 256-byte granules.
struct Arena {
  size_t size = 0, used = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
  void init(size_t n) {
    size = n & ~(size_t)255;
    free_.clear();
    if (size) free_[0] = size;
    used = 0;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t o = it->first + bytes;
      free_.erase(it);
      if (rest) free_[o] = rest;
      used += bytes;
      return true;
    }
    return false;
  }
  // Returns the free block the released range ends up in (after merging with
  // its neighbours) as (offset, bytes).
  std::pair<size_t, size_t> release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    used -= bytes;
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
        it = pv;
      }
    }
    return {it->first, it->second};
  }
  // Allocate `bytes` from the start of the free block at `block`.
  void take(size_t block, size_t bytes, size_t* off) {
    bytes = (bytes + 255) & ~(size_t)255;
    auto it = free_.find(block);
    const size_t rest = it->second - bytes;
    free_.erase(it);
    if (rest) free_[block + bytes] = rest;
    used += bytes;
    *off = block;
  }
};

}  // namespace pme
