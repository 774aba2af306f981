// Stress program for fddm::CopyPool (built and run by test_copy_pool.py):
// four threads copy random ranges at random offsets in 1..16 slices at once;
// every copied byte must match and the bytes around the range stay untouched.
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "fddm_copy_pool.hpp"

int main(int argc, char** argv) {
  const int iters = argc > 1 ? std::atoi(argv[1]) : 200;
  const size_t maxn = size_t(4) << 20, guard = 64;
  std::atomic<int> bad{0};
  std::vector<std::thread> callers;
  for (int c = 0; c < 4; ++c)
    callers.emplace_back([&, c] {
      std::mt19937_64 rng(1234 + c);
      std::vector<char> src(maxn + 2 * guard), dst(maxn + 2 * guard);
      for (size_t i = 0; i < src.size(); ++i) src[i] = static_cast<char>((i * 2654435761u >> 13) + c);
      for (int it = 0; it < iters; ++it) {
        const size_t off = guard + rng() % guard;
        size_t n = 1 + rng() % maxn;
        if (off + n > maxn + guard) n = maxn + guard - off;
        const unsigned parts = 1 + static_cast<unsigned>(rng() % 16);
        std::memset(dst.data(), 0x5a, dst.size());
        fddm::CopyPool::get().run(dst.data() + off, src.data() + off, n, parts);
        if (std::memcmp(dst.data() + off, src.data() + off, n) != 0) ++bad;
        for (size_t i = 0; i < off; ++i)
          if (dst[i] != 0x5a) { ++bad; break; }
        for (size_t i = off + n; i < dst.size(); ++i)
          if (dst[i] != 0x5a) { ++bad; break; }
      }
    });
  for (auto& t : callers) t.join();
  std::printf("copies %d, bad=%d\n", 4 * iters, bad.load());
  return bad.load() != 0;
}
