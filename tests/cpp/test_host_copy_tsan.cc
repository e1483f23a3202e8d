// The host copy pool (csrc/host_copy.cc) under ThreadSanitizer: four threads
// at once hand it jobs of pieces laid end to end (from a few bytes to tens of
// MB, so that some are copied on the calling thread and some spread over the
// pool), and every destination byte is checked.  Built and run by
// `make -C dials-metal_amd check-tsan`; exit status 0 when every copy is exact
// (a race ends the run through halt_on_error).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "host_copy.hh"

int main() {
  std::vector<int> bad(4, 0);
  std::vector<std::thread> ts;
  for (int t = 0; t < 4; t++)
    ts.emplace_back([t, &bad] {
      std::mt19937_64 rng(1000 + t);
      for (int job = 0; job < 12; job++) {
        const size_t pieces = 1 + rng() % 9;
        std::vector<std::vector<uint8_t>> src(pieces), dst(pieces);
        std::vector<sfi::CopyPiece> list;
        for (size_t i = 0; i < pieces; i++) {
          const size_t n = (job % 3 == 0) ? rng() % 4096 : rng() % (size_t(12) << 20);
          src[i].resize(n);
          dst[i].assign(n, 0);
          for (size_t k = 0; k < n; k += 4096)
            src[i][k] = static_cast<uint8_t>(rng());
          if (n)
            src[i][n - 1] = static_cast<uint8_t>(t + job);
          list.push_back(sfi::CopyPiece{dst[i].data(), src[i].data(), n});
        }
        sfi::host_copy(list.data(), list.size());
        for (size_t i = 0; i < pieces; i++)
          if (dst[i] != src[i])
            bad[t]++;
      }
    });
  for (std::thread &t : ts)
    t.join();
  int total = 0;
  for (int b : bad)
    total += b;
  std::printf("host copy pool, four threads: %s\n", total ? "FAILED" : "ok");
  return total ? 1 : 0;
}
