// glibc random_r TYPE_3 generator state and steps, usable from host and
// device code.  MathLibrary.cpp:27-33 (seed/randomReal) draws from it via
// srand/rand; keeping a per-box copy on the device lets batched boxes
// reproduce the reference's trial-move sequence without host round trips.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define MCGPU_HD __host__ __device__
#else
#define MCGPU_HD
#endif

namespace mcgpu {

// 31 state words, front index f and rear index r = f - 3 (mod 31).
// rand(): state[f] += state[r]; return the new word >> 1.
struct Rng31 {
  int32_t s[31];
  int32_t f;
  int32_t r;
};

MCGPU_HD inline int32_t rng31_next(Rng31& g) {
  const uint32_t v = (uint32_t)g.s[g.f] + (uint32_t)g.s[g.r];
  g.s[g.f] = (int32_t)v;
  g.f = (g.f == 30) ? 0 : g.f + 1;
  g.r = (g.r == 30) ? 0 : g.r + 1;
  return (int32_t)(v >> 1);
}

// srand(seed): Park-Miller words via Schrage's decomposition, 310 discards.
MCGPU_HD inline void rng31_seed(Rng31& g, uint32_t seed) {
  if (seed == 0) seed = 1;
  int32_t word = (int32_t)seed;
  g.s[0] = word;
  for (int i = 1; i < 31; i++) {
    const long long hi = word / 127773;
    const long long lo = word % 127773;
    long long w = 16807 * lo - 2836 * hi;
    if (w < 0) w += 2147483647;
    word = (int32_t)w;
    g.s[i] = word;
  }
  g.f = 3;
  g.r = 0;
  for (int i = 0; i < 310; i++) (void)rng31_next(g);
}

// MathLibrary.cpp:31-33: (end - start) * ((Real)rand() / RAND_MAX) + start
MCGPU_HD inline double random_real(int32_t r, double start, double end) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  return (end - start) * ((double)r / 2147483647) + start;
}

}  // namespace mcgpu
