// ya_philox.hpp -- the random stream of the samplers: Philox4x64-10
// (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC 2011), counter-based, so any block of the stream is computed on its own
// and host and device produce the same words.
//
// The reference draws its samplers' randomness from an unseeded Salsa20
// keystream (lib/prng/fastrandombytes.cpp:21-34), which no seed reproduces.
// Here a sampling call with seed s reads the byte stream
//     B(s) = block(0) || block(1) || ...,
//     block(b) = Philox4x64-10(counter = (b, 0, 0, 0), key = (s, 0))
// as four little-endian 64-bit words (32 bytes), and the hwt sampler's
// reservoir step k reads its own block, counter (k, 1, 0, 0).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ya {

struct Philox4 {
  uint64_t v[4];
};

// 64 x 64 -> 128-bit product as (hi, lo).
__host__ __device__ __forceinline__ void philox_mulhilo(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
  const unsigned __int128 p = (unsigned __int128)a * b;
  hi = (uint64_t)(p >> 64);
  lo = (uint64_t)p;
}

// Ten rounds; the round multipliers and the key increments (Weyl sequence of
// the golden ratio and of sqrt(3) - 1) are those of the paper's Philox4x64.
__host__ __device__ __forceinline__ Philox4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                                                          uint64_t k0, uint64_t k1) {
  constexpr uint64_t kM0 = 0xD2E7470EE14C6C93ULL, kM1 = 0xCA5A826395121157ULL;
  constexpr uint64_t kW0 = 0x9E3779B97F4A7C15ULL, kW1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += kW0;
      k1 += kW1;
    }
    uint64_t h0, l0, h1, l1;
    philox_mulhilo(kM0, c0, h0, l0);
    philox_mulhilo(kM1, c2, h1, l1);
    const uint64_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
    c0 = n0;
    c1 = l1;
    c2 = n2;
    c3 = l0;
  }
  return Philox4{{c0, c1, c2, c3}};
}

// Block b of the main stream of seed s.
__host__ __device__ __forceinline__ Philox4 stream_block(uint64_t seed, uint64_t b) {
  return philox4x64_10(b, 0, 0, 0, seed, 0);
}

// 64-bit word j of the main stream (selects rather than a variable index,
// which on the device would move the block out of registers).
__host__ __device__ __forceinline__ uint64_t stream_word(uint64_t seed, uint64_t j) {
  const Philox4 r = stream_block(seed, j >> 2);
  const uint32_t k = (uint32_t)(j & 3);
  const uint64_t lo = (k & 1) ? r.v[1] : r.v[0], hi = (k & 1) ? r.v[3] : r.v[2];
  return (k & 2) ? hi : lo;
}

// Reservoir step k of the hwt sampler: its own block.
__host__ __device__ __forceinline__ Philox4 reservoir_block(uint64_t seed, uint64_t k) {
  return philox4x64_10(k, 1, 0, 0, seed, 0);
}

// poly::set(hwt_dist) (include/yell/details/poly.hpp:296-315): a position
// drawn uniformly below k by rejection -- a random word w is accepted when
// w <= floor(2^64-1 / k) * k, then reduced mod k.  Here step k tries the four
// words of its block in order; if all four are rejected (probability below
// (k / 2^64)^4) the last one is reduced anyway.
__host__ __device__ __forceinline__ uint64_t reservoir_pos(uint64_t seed, uint64_t k) {
  const Philox4 r = reservoir_block(seed, k);
  const uint64_t lim = (~0ULL / k) * k;
  uint64_t w = r.v[3];
  if (r.v[2] <= lim) w = r.v[2];
  if (r.v[1] <= lim) w = r.v[1];
  if (r.v[0] <= lim) w = r.v[0];
  return w % k;
}

}  // namespace ya
