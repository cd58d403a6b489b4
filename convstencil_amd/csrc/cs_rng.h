// cs_rng.h -- the synthetic-input generator, shared by host and device code.
//
// SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014).  Its r-th output from state s0 is mix(s0 + (r+1)*G),
// so the device fills element r independently and produces the same bits as a
// sequential host loop.
//   stream seed : one SplitMix64 draw from (seed * 2 + stream), stream 0 = grid
//                 data, 1 = weights
//   grid data   : interior cells in row-major interior order, (x >> 11) * 2^-53 * 2 - 1
//                 in [-1, 1); every halo cell 0
//   weights     : ((x >> 11) + 0.5) * 2^-53 in (0, 1), divided by their
//                 left-to-right double sum
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define CS_HD __host__ __device__
#else
#define CS_HD
#endif

namespace cs {

static constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

CS_HD inline uint64_t splitmix_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

CS_HD inline uint64_t splitmix_next(uint64_t &state) {
  state += kGolden;
  return splitmix_mix(state);
}

CS_HD inline uint64_t stream_seed(uint64_t seed, uint64_t stream) {
  uint64_t s = seed * 2ULL + stream;
  return splitmix_next(s);
}

// r-th draw (0-based) of the stream that starts at state s0
CS_HD inline uint64_t splitmix_at(uint64_t s0, uint64_t r) { return splitmix_mix(s0 + (r + 1) * kGolden); }

CS_HD inline double unit_data(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53 * 2.0 - 1.0; }
CS_HD inline double unit_weight(uint64_t x) { return ((double)(x >> 11) + 0.5) * 0x1.0p-53; }

}  // namespace cs
