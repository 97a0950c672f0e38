// Seeded pseudo-random streams for the synthetic hypergraph generator.
//
// splitmix64 expands a 64-bit seed into the 256-bit state of xoshiro256**
// (Blackman & Vigna, "Scrambled linear pseudorandom number generators", 2018).
// Both are pure integer recurrences, so a seed produces the same stream on
// every machine; doubles are formed from the top 53 bits.
#pragma once
#include <cstdint>
#include <cmath>

namespace hm {

inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // Uniform double in [0, 1).
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  // Uniform integer in [0, n), n > 0 (multiply-shift; bias below 2^-32 for n < 2^32).
  uint64_t below(uint64_t n) {
    unsigned __int128 p = (unsigned __int128)next() * n;
    return (uint64_t)(p >> 64);
  }
  // Exponential variate with the given mean.
  double exponential(double mean) { return -mean * std::log1p(-uniform()); }
};

}  // namespace hm
