// arena.hpp -- the server's device-memory arena: a first-fit allocator of
// byte offsets in [0, size), in 64 KiB granules, with coalescing of freed
// ranges.  Plain C++ with no HIP dependency (built under the sanitizers by
// tests/native).
#pragma once

#include <cstddef>
#include <iterator>
#include <map>

namespace siting {

class Arena {
 public:
  void reset(size_t size) {
    free_.clear();
    if (size) free_[0] = size;
    size_ = size;
  }
  static size_t granule(size_t b) { return (b + 65535) / 65536 * 65536; }
  bool alloc(size_t bytes, size_t *off) {
    bytes = granule(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = granule(bytes);
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
  }
  size_t size() const { return size_; }
  size_t free_bytes() const {
    size_t t = 0;
    for (const auto &kv : free_) t += kv.second;
    return t;
  }
  size_t free_ranges() const { return free_.size(); }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0;
};

}  // namespace siting
