// Deterministic restatement of GNU libstdc++'s std::shuffle driven by
// std::default_random_engine (minstd_rand0), the leaf shuffle of
// SamplerRandom (Converter/include/sampler_random.h:97-117).
//
// minstd_rand0 is the Park-Miller generator x' = 16807 x mod (2^31 - 1);
// seeding with s sets x = s mod (2^31 - 1), or 1 when that is 0, and every
// call returns the next state (range [1, 2^31 - 2]).
//
// libstdc++'s shuffle is Fisher-Yates from the front.  When the generator's
// range (max - min = 2^31 - 3) divided by n is at least n, it draws two
// swap positions from one number: an even n first swaps element 1 with
// element d{0,1}; then for i = 1 or 2, 3, ... in pairs, x = d{0, (i+1)(i+2) - 1}
// gives positions x / (i + 2) for element i and x % (i + 2) for element i + 1.
// Otherwise element i swaps with d{0, i} for i = 1 .. n-1.  d{a, b} is
// uniform_int_distribution's downscaling: with range r = b - a + 1,
// scaling = (2^31 - 3) / r and past = r * scaling, draw u = g() - 1 until
// u < past, then a + u / scaling.  The generator's range is neither 2^32 - 1
// nor 2^64 - 1, so no other branch of uniform_int_distribution applies.
// Host and device (single lane) use the same code.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define LSHUF_HD __host__ __device__ inline
#else
#define LSHUF_HD inline
#endif

namespace lstd {

struct MinStd0 {
  uint32_t x;
  LSHUF_HD explicit MinStd0(uint32_t seed) {
    x = seed % 2147483647u;
    if (x == 0) x = 1;
  }
  // 16807 x mod (2^31 - 1) with the Mersenne-prime fold
  LSHUF_HD uint32_t operator()() {
    const uint64_t t = (uint64_t)x * 16807u;
    uint32_t r = (uint32_t)(t & 0x7FFFFFFFu) + (uint32_t)(t >> 31);
    if (r >= 2147483647u) r -= 2147483647u;
    x = r;
    return r;
  }
};

constexpr uint32_t kMinStdRange = 2147483645u;  // max() - min()

// uniform_int_distribution<>{0, r - 1}(g) for 1 <= r <= kMinStdRange
LSHUF_HD uint32_t uniform_below(MinStd0& g, uint32_t r) {
  const uint32_t scaling = kMinStdRange / r;
  const uint32_t past = r * scaling;
  uint32_t u;
  do {
    u = g() - 1u;
  } while (u >= past);
  return u / scaling;
}

// Calls swap(i, j) in the order std::shuffle swaps a[i] with a[j] on an
// array of n elements with std::default_random_engine(seed).
template <class Swap>
LSHUF_HD void shuffle_swaps(int64_t n, uint32_t seed, Swap swap) {
  if (n <= 1) {
    // std::shuffle draws nothing for an empty range and, for one element,
    // nothing either (both loops are empty)
    return;
  }
  MinStd0 g(seed);
  const uint64_t un = (uint64_t)n;
  if ((uint64_t)kMinStdRange / un >= un) {
    int64_t i = 1;
    if ((un % 2) == 0) {
      swap(i, (int64_t)uniform_below(g, 2));
      i++;
    }
    while (i != n) {
      const uint32_t b0 = (uint32_t)i + 1u;  // swap range of element i
      const uint32_t b1 = b0 + 1u;
      const uint32_t x = uniform_below(g, b0 * b1);
      swap(i, (int64_t)(x / b1));
      i++;
      swap(i, (int64_t)(x % b1));
      i++;
    }
    return;
  }
  for (int64_t i = 1; i != n; ++i) swap(i, (int64_t)uniform_below(g, (uint32_t)i + 1u));
}

}  // namespace lstd
