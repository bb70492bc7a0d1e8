// yell/prng/seed.hpp -- the seed of the drop-in's samplers.
//
// The reference keys its samplers from /dev/urandom
// (lib/prng/fastrandombytes.cpp:21-34), so its inputs cannot be reproduced.
// Here every sampling call (poly::set, FastGaussianNoise::getNoise) takes the
// next value of one seed counter, so a program's inputs are reproducible:
// set_seed(s) restarts the sequence.
#pragma once
#include <atomic>
#include <cstdint>

namespace yell {

namespace detail {
inline std::atomic<uint64_t> &seed_counter() {
  static std::atomic<uint64_t> s{0x5eed5eed5eedULL};
  return s;
}
}  // namespace detail

inline void set_seed(uint64_t s) { detail::seed_counter().store(s); }
// Returns the current seed and advances the counter (by the 64-bit golden
// ratio, so consecutive seeds differ in many bits).
inline uint64_t next_seed() { return detail::seed_counter().fetch_add(0x9E3779B97F4A7C15ULL); }

}  // namespace yell
