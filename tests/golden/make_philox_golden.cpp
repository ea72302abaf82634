// Writes philox_rocrand.txt: Philox4x32-10 outputs for random (counter, key)
// pairs, computed by rocRAND's philox4x32_10_engine (an implementation
// independent of this project's). Build and run:
//   hipcc -O1 -I/opt/rocm/include make_philox_golden.cpp -o /tmp/mkphilox && /tmp/mkphilox > philox_rocrand.txt
// Mapping onto the engine: key = seed, counter words 2..3 = subsequence,
// counter words 0..1 = offset / 4 (rocRAND restart()/discard semantics).
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

static uint64_t sm = 0x9E3779B97F4A7C15ull;
static uint64_t splitmix() {
  uint64_t z = (sm += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int main() {
  printf("# c0 c1 c2 c3 k0 k1 -> o0 o1 o2 o3 (hex), from rocRAND philox4x32_10_engine\n");
  for (int t = 0; t < 64; ++t) {
    uint64_t a = splitmix(), b = splitmix(), k = splitmix();
    a &= (1ull << 62) - 1;  // the engine takes the offset as 4 * (counter words 0..1)
    if (t < 8) { a &= 0xffffull; b &= 0xffull; }  // small counters, like chain/step indices
    uint32_t c0 = (uint32_t)a, c1 = (uint32_t)(a >> 32), c2 = (uint32_t)b, c3 = (uint32_t)(b >> 32);
    rocrand_device::philox4x32_10_engine e(k, b, 4ull * a);
    uint4 r = e.next4();
    printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", c0, c1, c2, c3, (uint32_t)k,
           (uint32_t)(k >> 32), r.x, r.y, r.z, r.w);
  }
  return 0;
}
