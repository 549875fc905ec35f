// CPU test of the server's staging copy (glscv_amd/csrc/host_copy.hpp): the
// parts cover every byte exactly once, for byte counts that are and are not
// multiples of the part count and of 64.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../glscv_amd/csrc/host_copy.hpp"

static int check_cover(size_t bytes, int parts) {
  std::vector<int> hits(bytes, 0);
  for (int k = 0; k < parts; ++k) {
    size_t off = 0, n = 0;
    glsx::copy_part(bytes, parts, k, &off, &n);
    if (off + n > bytes) return 1;
    if (n && k && off % 64) return 2;
    for (size_t i = off; i < off + n; ++i) hits[i]++;
  }
  for (size_t i = 0; i < bytes; ++i)
    if (hits[i] != 1) return 3;
  return 0;
}

int main() {
  const size_t sizes[] = {1, 63, 64, 65, 191, 192, 256, 1000, 4096 + 3, 12 * 64 + 1, (size_t)1 << 20, 3 * 65 * 64 + 7};
  for (size_t b : sizes)
    for (int parts = 1; parts <= 5; ++parts)
      if (int rc = check_cover(b, parts)) {
        std::printf("FAIL cover bytes=%zu parts=%d rc=%d\n", b, parts, rc);
        return 1;
      }
  // the copy itself, split into 3 and 4 threads with a small split size: an
  // odd byte count whose per-part share is a multiple of 64 (3 * 64 * 5 + 1)
  const size_t odd[] = {3 * 64 * 5 + 1, 4 * 64 * 7 + 3, 1234567};
  for (size_t b : odd)
    for (int th = 2; th <= 4; ++th) {
      std::vector<unsigned char> src(b), dst(b, 0);
      for (size_t i = 0; i < b; ++i) src[i] = (unsigned char)(i * 131 + 7);
      glsx::staging_copy(dst.data(), src.data(), b, th, 64);
      for (size_t i = 0; i < b; ++i)
        if (dst[i] != src[i]) {
          std::printf("FAIL copy bytes=%zu threads=%d first bad byte %zu\n", b, th, i);
          return 1;
        }
    }
  std::printf("host copy ok\n");
  return 0;
}
