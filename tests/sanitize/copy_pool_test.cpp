// Exercises ustbx::CopyPool (ustb_amd/csrc/host_copy.h), the host-buffer
// entry's threaded memcpy, under the host sanitizers: copies of sizes around
// the pool's 1 MiB threshold and its 64-byte part boundaries, from two caller
// threads at once, each checked byte for byte; exit status 0 when all match.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <thread>
#include <vector>

#include "host_copy.h"

static int check(ustbx::CopyPool& pool, unsigned seed, int rounds) {
  const size_t sizes[] = {0, 1, 63, 64, 4095, (1u << 20) - 1, 1u << 20, (1u << 20) + 1, (3u << 20) + 17,
                          (8u << 20), (8u << 20) + 333};
  std::vector<unsigned char> src(sizeof(size_t) + (9u << 20)), dst(src.size());
  for (size_t i = 0; i < src.size(); ++i) src[i] = (unsigned char)((i * 2654435761u + seed) >> 13);
  int bad = 0;
  for (int r = 0; r < rounds; ++r) {
    for (size_t n : sizes) {
      const size_t off = (size_t)((r * 7 + seed) % 5);  // unaligned starts too
      memset(dst.data(), 0xA5, dst.size());
      pool.copy(dst.data() + off, src.data() + off, n);
      if (memcmp(dst.data() + off, src.data() + off, n) != 0) ++bad;
      // bytes outside the range untouched
      if (off && dst[off - 1] != 0xA5) ++bad;
      if (off + n < dst.size() && dst[off + n] != 0xA5) ++bad;
    }
  }
  return bad;
}

int main() {
  for (int threads : {1, 2, 4, 7}) {
    ustbx::CopyPool pool(threads);
    int bad1 = 0, bad2 = 0;
    std::thread a([&] { bad1 = check(pool, 1, 3); });
    std::thread b([&] { bad2 = check(pool, 2, 3); });
    a.join();
    b.join();
    if (bad1 || bad2) {
      printf("threads %d: %d mismatches\n", threads, bad1 + bad2);
      return 1;
    }
  }
  printf("ok\n");
  return 0;
}
