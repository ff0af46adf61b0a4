// CPU check of cdn::CopyPool (cudadecon_amd/csrc/copypool.hpp): every byte
// of every job lands, for sizes around the 2 MB slice minimum and ragged
// tails, jobs with fewer slices than workers, back-to-back jobs, and pools
// created and destroyed.  Built with ThreadSanitizer by tests/test_copypool.py.
#include "copypool.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint8_t next_byte() {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return (uint8_t)rng;
}

int main() {
  const size_t MB = 1UL << 20;
  const size_t sizes[] = {0, 1, 2 * MB - 1, 2 * MB, 4 * MB + 3, 6 * MB + 1, 15 * MB + 5, 33 * MB + 7};
  std::vector<uint8_t> src(34 * MB), dst(34 * MB + 64);
  for (auto& b : src) b = next_byte();
  int fails = 0;
  for (int round = 0; round < 2; ++round) {
    cdn::CopyPool pool;
    for (int rep = 0; rep < 3; ++rep)
      for (size_t n : sizes)
        for (size_t off : {(size_t)0, (size_t)3}) {
          std::memset(dst.data(), 0xA5, dst.size());
          pool.copy(dst.data() + off, src.data() + off, n);
          const bool body = std::memcmp(dst.data() + off, src.data() + off, n) == 0;
          bool guard = true;
          for (size_t i = 0; i < off; ++i) guard = guard && dst[i] == 0xA5;
          for (size_t i = off + n; i < off + n + 64 && i < dst.size(); ++i) guard = guard && dst[i] == 0xA5;
          if (!body || !guard) {
            std::printf("FAIL n=%zu off=%zu body=%d guard=%d\n", n, off, (int)body, (int)guard);
            ++fails;
          }
        }
    std::printf("pool of %zu threads: %s\n", pool.threads(), fails ? "FAIL" : "ok");
  }
  return fails ? 1 : 0;
}
