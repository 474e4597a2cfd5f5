// Cross-check generator: Philox4x32-10 blocks from the open-source rocRAND
// engine (host path), printed as JSON.  Counter (x, y, z, w) and key (k0, k1)
// map to the engine as seed = k0 | k1 << 32, subsequence = z | w << 32,
// offset = 4 * (x | y << 32), applied in pieces that do not overflow.
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

int main() {
  const uint32_t cases[][6] = {
      {0, 0, 0, 0, 0, 0},
      {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
      {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
      {0, 0, 0, 0, 7, 0},
      {5, 3, 1, 0, 12345, 2},
      {4999, 4095, 1, 0, 42, 0},
      {17, 123, 2, 3, 0xdeadbeefu, 7},
  };
  printf("[\n");
  const int n = sizeof(cases) / sizeof(cases[0]);
  for (int i = 0; i < n; ++i) {
    const uint32_t* c = cases[i];
    unsigned long long seed = (unsigned long long)c[4] | ((unsigned long long)c[5] << 32);
    unsigned long long sub = (unsigned long long)c[2] | ((unsigned long long)c[3] << 32);
    unsigned long long blocks = (unsigned long long)c[0] | ((unsigned long long)c[1] << 32);
    rocrand_device::philox4x32_10_engine eng(seed, sub, 0);
    while (blocks) {
      unsigned long long step = blocks < (1ull << 61) ? blocks : (1ull << 61);
      eng.discard(4ull * step);
      blocks -= step;
    }
    uint32_t o[4];
    for (int k = 0; k < 4; ++k) o[k] = eng.next();
    printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
           c[0], c[1], c[2], c[3], c[4], c[5], o[0], o[1], o[2], o[3], i + 1 < n ? "," : "");
  }
  printf("]\n");
  return 0;
}
