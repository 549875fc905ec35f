// host_copy.hpp -- the server's copy of a submission into pinned staging,
// split over a few host threads for large images (plain C++, no HIP: the CPU
// tests compile it with g++).
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <thread>
#include <vector>

namespace glsx {

// One thread's memcpy (about 20 GB/s) bounds a burst of submissions below what
// the uploads and the device sustain, so images of kCopySplit bytes and more
// are copied by up to kCopyThreads threads (the serving mix of 256^2..2048^2
// pairs: a burst of 2,670 req/s with 4 threads of 4 MiB or more, 3,080 with 8
// of 2 MiB or more; 16 threads measured 2,926-2,943 against 2,975-3,013 for 8
// on one box).
constexpr int kCopyThreads = 8;
constexpr size_t kCopySplit = (size_t)2 << 20;  // bytes per thread at least

// Part k of `parts` of a copy of `bytes`: [*off, *off + *n).  Parts start on
// 64-byte boundaries; the last part runs to the end, so the parts cover every
// byte for any `bytes` (parts with nothing left get n = 0).
inline void copy_part(size_t bytes, int parts, int k, size_t* off, size_t* n) {
  const size_t chunk = ((bytes + (size_t)parts - 1) / (size_t)parts + 63) & ~(size_t)63;
  const size_t o = std::min(bytes, (size_t)k * chunk);
  const size_t end = k == parts - 1 ? bytes : std::min(bytes, o + chunk);
  *off = o;
  *n = end - o;
}

inline void staging_copy(void* dst, const void* src, size_t bytes, int max_threads = kCopyThreads,
                         size_t split = kCopySplit) {
  const int parts = (int)std::min<size_t>((size_t)std::max(1, max_threads), std::max<size_t>(1, bytes / split));
  if (parts <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> helpers;
  helpers.reserve(parts - 1);
  for (int k = 1; k < parts; ++k) {
    size_t off = 0, n = 0;
    copy_part(bytes, parts, k, &off, &n);
    if (n == 0) continue;
    try {
      helpers.emplace_back([=] { std::memcpy((char*)dst + off, (const char*)src + off, n); });
    } catch (...) {  // no thread available: copy this part here
      std::memcpy((char*)dst + off, (const char*)src + off, n);
    }
  }
  size_t off0 = 0, n0 = 0;
  copy_part(bytes, parts, 0, &off0, &n0);
  std::memcpy(dst, src, n0);
  for (auto& h : helpers) h.join();
}

}  // namespace glsx
