// ya_sampler.cpp -- seeded host samplers, the input source of the engine.
//
// The reference's samplers draw from an unseeded Salsa20 keystream
// (lib/prng/fastrandombytes.cpp:21-34), so their outputs cannot be
// reproduced; these keep the reference's mapping of random words to residues
// and are seeded:
//   uniform   include/yell/details/poly.hpp:372-396
//   gaussian  include/yell/details/poly.hpp:408-431 (one signed sample per
//             coefficient shared by all residues, r < 0 ? p + r : r)
// The random words come from xoshiro256** (Blackman & Vigna) seeded through
// splitmix64; the Gaussian is a cumulative-distribution-table sampler over
// |x| <= ceil(13 sigma) (tail mass below 2^-120).
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <vector>

#include "../../include/yell_amd.h"
#include "ya_internal.hpp"

namespace {

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed) {
    for (auto &w : s) {  // splitmix64
      seed += 0x9e3779b97f4a7c15ULL;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      w = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
};

}  // namespace

namespace {

// Shape check of either build: degree a power of two in [8, 2^maxlog],
// nmoduli in [1, nmax].
int sample_shape(size_t degree, size_t nmoduli, int maxlog, size_t nmax) {
  if (degree < 8 || (degree & (degree - 1)) || degree > ((size_t)1 << maxlog))
    return ya::fail(YA_ERR_INVALID, "degree %zu: a power of two in [8, 2^%d] is required", degree, maxlog);
  if (nmoduli == 0 || nmoduli > nmax) return ya::fail(YA_ERR_INVALID, "nmoduli %zu out of [1, %zu]", nmoduli, nmax);
  return YA_OK;
}

// include/yell/details/poly.hpp:372-396: per residue, a random word masked to
// the prime's bit length (kModulusBitsize), minus p once if >= p.
template <class W, class PF>
int sample_uniform(W *out, size_t degree, size_t nmoduli, uint64_t seed, PF prime) {
  if (!out) return ya::fail(YA_ERR_INVALID, "null output");
  Xoshiro g(seed);
  for (size_t cm = 0; cm < nmoduli; ++cm) {
    const uint64_t p = prime(cm);
    const int top = 63 - __builtin_clzll(p);  // floor(log2 p)
    const uint64_t mask = (top >= 63) ? ~0ULL : ((1ULL << (top + 1)) - 1);
    W *dst = out + cm * degree;
    for (size_t i = 0; i < degree; ++i) {
      uint64_t v = g.next() & mask;
      if (v >= p) v -= p;
      dst[i] = (W)v;
    }
  }
  return YA_OK;
}

template <class W, class PF>
int sample_gaussian(W *out, size_t degree, size_t nmoduli, double sigma, uint64_t seed, PF prime) {
  if (!out) return ya::fail(YA_ERR_INVALID, "null output");
  if (!(sigma > 0.0) || sigma > 1e6) return ya::fail(YA_ERR_INVALID, "sigma %g out of range", sigma);
  const int tail = (int)std::ceil(13.0 * sigma);
  std::vector<long double> w(tail + 1);
  long double tot = 0;
  for (int k = 0; k <= tail; ++k) {
    w[k] = std::exp(-(long double)k * k / (2.0L * sigma * sigma)) * (k ? 2.0L : 1.0L);
    tot += w[k];
  }
  // cdf[k] = P(|X| <= k) scaled to 2^64; the last entry saturates.
  std::vector<uint64_t> cdf(tail + 1);
  long double acc = 0;
  for (int k = 0; k <= tail; ++k) {
    acc += w[k] / tot;
    const long double v = acc * 18446744073709551616.0L;
    cdf[k] = (k == tail || v >= 18446744073709551615.0L) ? ~0ULL : (uint64_t)v;
  }
  Xoshiro g(seed);
  std::vector<int64_t> r(degree);
  for (size_t i = 0; i < degree; ++i) {
    const uint64_t u = g.next();
    int k = 0;
    while (k < tail && u >= cdf[k]) ++k;
    const bool neg = k != 0 && (g.next() >> 63);
    r[i] = neg ? -(int64_t)k : (int64_t)k;
  }
  for (size_t cm = 0; cm < nmoduli; ++cm) {
    const int64_t p = (int64_t)prime(cm);
    W *dst = out + cm * degree;
    for (size_t i = 0; i < degree; ++i) {
      // reduce |r| < p first (sigma large relative to a 30-bit prime), then map r < 0 to p + r
      const int64_t v = r[i] % p;
      dst[i] = (W)(v < 0 ? p + v : v);
    }
  }
  return YA_OK;
}

// poly::set(ZO_dist) (include/yell/details/poly.hpp:438-451): one random byte
// r per coefficient shared by all residues; r <= rho ? ((r & 2) ? p-1 : 1) : 0.
template <class W, class PF>
int sample_zo(W *out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed, PF prime) {
  if (!out || rho > 255) return ya::fail(YA_ERR_INVALID, "null output or rho > 255");
  Xoshiro g(seed);
  std::vector<uint8_t> r(degree);
  for (size_t i = 0; i < degree; i += 8) {
    const uint64_t w = g.next();
    for (size_t k = 0; k < 8 && i + k < degree; ++k) r[i + k] = (uint8_t)(w >> (8 * k));
  }
  for (size_t cm = 0; cm < nmoduli; ++cm) {
    const W p = (W)prime(cm);
    for (size_t i = 0; i < degree; ++i)
      out[cm * degree + i] = r[i] <= rho ? ((r[i] & 2) ? p - 1 : 1) : 0;
  }
  return YA_OK;
}

// poly::set(hwt_dist) (include/yell/details/poly.hpp:458-501): exactly `hwt`
// positions chosen uniformly (reservoir sampling), each +1 or -1 with equal
// probability, shared by all residues (-1 stored as p-1).
template <class W, class PF>
int sample_hwt(W *out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed, PF prime) {
  if (!out || hwt == 0 || hwt > degree) return ya::fail(YA_ERR_INVALID, "hwt must be in [1, degree]");
  Xoshiro g(seed);
  std::vector<size_t> hit(hwt);
  for (size_t k = 0; k < hwt; ++k) hit[k] = k;
  for (size_t k = hwt; k < degree; ++k) {  // reservoir sampling, unbiased draw in [0, k]
    const uint64_t lim = UINT64_MAX - UINT64_MAX % (k + 1);
    uint64_t v;
    do v = g.next(); while (v >= lim);
    const size_t pos = (size_t)(v % (k + 1));
    if (pos < hwt) hit[pos] = k;
  }
  std::vector<uint8_t> sign(hwt);
  for (size_t k = 0; k < hwt; ++k) sign[k] = (uint8_t)(g.next() >> 63);
  for (size_t cm = 0; cm < nmoduli; ++cm) {
    const W p = (W)prime(cm);
    W *dst = out + cm * degree;
    std::fill(dst, dst + degree, 0);
    for (size_t k = 0; k < hwt; ++k) dst[hit[k]] = sign[k] ? 1 : p - 1;
  }
  return YA_OK;
}

uint64_t prime64(size_t cm) { return ya::prime(cm); }
uint64_t prime32(size_t cm) { return ya32_prime(cm); }

int shape64(size_t degree, size_t nmoduli) {
  int logn;
  return ya::check_shape(degree, 0, nmoduli, &logn);
}
int shape32(size_t degree, size_t nmoduli) { return sample_shape(degree, nmoduli, 15, ya32_num_moduli()); }

}  // namespace

extern "C" {

int ya_sample_uniform(uint64_t *out, size_t degree, size_t nmoduli, uint64_t seed) {
  int rc = shape64(degree, nmoduli);
  return rc ? rc : sample_uniform(out, degree, nmoduli, seed, prime64);
}
int ya_sample_gaussian(uint64_t *out, size_t degree, size_t nmoduli, double sigma, uint64_t seed) {
  int rc = shape64(degree, nmoduli);
  return rc ? rc : sample_gaussian(out, degree, nmoduli, sigma, seed, prime64);
}
int ya_sample_zo(uint64_t *out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed) {
  int rc = shape64(degree, nmoduli);
  return rc ? rc : sample_zo(out, degree, nmoduli, rho, seed, prime64);
}
int ya_sample_hwt(uint64_t *out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed) {
  int rc = shape64(degree, nmoduli);
  return rc ? rc : sample_hwt(out, degree, nmoduli, hwt, seed, prime64);
}

int ya32_sample_uniform(uint32_t *out, size_t degree, size_t nmoduli, uint64_t seed) {
  int rc = shape32(degree, nmoduli);
  return rc ? rc : sample_uniform(out, degree, nmoduli, seed, prime32);
}
int ya32_sample_gaussian(uint32_t *out, size_t degree, size_t nmoduli, double sigma, uint64_t seed) {
  int rc = shape32(degree, nmoduli);
  return rc ? rc : sample_gaussian(out, degree, nmoduli, sigma, seed, prime32);
}
int ya32_sample_zo(uint32_t *out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed) {
  int rc = shape32(degree, nmoduli);
  return rc ? rc : sample_zo(out, degree, nmoduli, rho, seed, prime32);
}
int ya32_sample_hwt(uint32_t *out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed) {
  int rc = shape32(degree, nmoduli);
  return rc ? rc : sample_hwt(out, degree, nmoduli, hwt, seed, prime32);
}

}  // extern "C"
