// glibc_rand.hpp -- the first value rand() returns after srand(seed) in the
// GNU C library (its default TYPE_3 additive feedback generator).  The
// reference draws its ray-casting directions with srand(n); rand()
// (ibm_iso.c:69-71), so the classification is only reproducible with the
// same sequence.  Restated from the generator's published definition
// (random_r.c: a Park-Miller LCG seeds r[0..30], r[i] = r[i-31] + r[i-3]
// mod 2^32 thereafter, the first 310 outputs are discarded, and rand()
// returns r[i] >> 1); host code only.
#pragma once
#include <cstdint>

namespace cles {

inline int32_t glibc_first_rand(uint32_t seed) {
  if (seed == 0) seed = 1;  // srand(0) seeds like srand(1)
  uint32_t r[344 + 1];
  int32_t w = (int32_t)seed;
  r[0] = (uint32_t)w;
  for (int i = 1; i < 31; i++) {
    // 16807 * r[i-1] mod (2^31 - 1) with Schrage's decomposition
    const int32_t hi = w / 127773, lo = w % 127773;
    w = 16807 * lo - 2836 * hi;
    if (w < 0) w += 2147483647;
    r[i] = (uint32_t)w;
  }
  for (int i = 31; i < 34; i++) r[i] = r[i - 31];
  for (int i = 34; i <= 344; i++) r[i] = r[i - 31] + r[i - 3];
  return (int32_t)(r[344] >> 1);
}

}  // namespace cles
