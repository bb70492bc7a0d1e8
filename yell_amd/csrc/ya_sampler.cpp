// ya_sampler.cpp -- seeded host samplers, the input source of the engine.
//
// The reference's samplers draw from an unseeded Salsa20 keystream
// (lib/prng/fastrandombytes.cpp:21-34), so their outputs cannot be
// reproduced.  These keep the reference's mapping of random bytes to
// residues and read the seeded Philox stream of ya_philox.hpp; the device
// samplers (ya_sampler_dev.hip) read the same stream and return the same
// words:
//   uniform   include/yell/details/poly.hpp:200-228
//   gaussian  include/yell/details/poly.hpp:236-260 (one signed sample per
//             coefficient shared by all residues, r < 0 ? p + r : r)
//   ZO        include/yell/details/poly.hpp:267-279
//   hwt       include/yell/details/poly.hpp:287-330
// The Gaussian is a cumulative-distribution-table sampler over
// |x| <= ceil(13 sigma) (tail mass below 2^-120); the reference uses its own
// table sampler (include/yell/prng/FastGaussianNoise.hpp:484-602) for the same
// distribution.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/yell_amd.h"
#include "ya_internal.hpp"
#include "ya_philox.hpp"

namespace ya {

// cdf[k] = P(|X| <= k) * 2^64 for k < tail, cdf[tail] = 2^64 - 1, with
// P(X = x) proportional to exp(-x^2 / (2 sigma^2)) over |x| <= tail.
int gaussian_cdt(double sigma, std::vector<uint64_t> &cdf) {
  if (!(sigma > 0.0) || sigma > kMaxGaussianSigma)
    return fail(YA_ERR_INVALID, "sigma %g out of (0, %g]", sigma, kMaxGaussianSigma);
  const int tail = (int)std::ceil(13.0 * sigma);
  std::vector<long double> w(tail + 1);
  long double tot = 0;
  for (int k = 0; k <= tail; ++k) {
    w[k] = std::exp(-(long double)k * k / (2.0L * sigma * sigma)) * (k ? 2.0L : 1.0L);
    tot += w[k];
  }
  cdf.assign(tail + 1, 0);
  long double acc = 0;
  for (int k = 0; k <= tail; ++k) {
    acc += w[k] / tot;
    const long double v = acc * 18446744073709551616.0L;
    cdf[k] = (k == tail || v >= 18446744073709551615.0L) ? ~0ULL : (uint64_t)v;
  }
  return YA_OK;
}

}  // namespace ya

namespace {

// Shape check of either build: degree a power of two in [8, 2^maxlog],
// nmoduli in [1, nmax].
int sample_shape(size_t degree, size_t nmoduli, int maxlog, size_t nmax) {
  if (degree < 8 || (degree & (degree - 1)) || degree > ((size_t)1 << maxlog))
    return ya::fail(YA_ERR_INVALID, "degree %zu: a power of two in [8, 2^%d] is required", degree, maxlog);
  if (nmoduli == 0 || nmoduli > nmax) return ya::fail(YA_ERR_INVALID, "nmoduli %zu out of [1, %zu]", nmoduli, nmax);
  return YA_OK;
}

// Word g (of type W) of the byte stream of `seed`.
template <class W>
W stream_value(uint64_t seed, size_t g) {
  if constexpr (sizeof(W) == 8) return ya::stream_word(seed, g);
  return (W)(ya::stream_word(seed, g >> 1) >> (32 * (g & 1)));
}

// poly::set(uniform) (include/yell/details/poly.hpp:200-228): residue cm
// takes the next degree words of the stream; each is masked to the prime's
// bit length (floor(log2 p) + 1 bits) and reduced by p once if >= p.
template <class W, class PF>
int sample_uniform(W *out, size_t degree, size_t nmoduli, uint64_t seed, PF prime) {
  if (!out) return ya::fail(YA_ERR_INVALID, "null output");
  for (size_t cm = 0; cm < nmoduli; ++cm) {
    const uint64_t p = prime(cm);
    const int top = 63 - __builtin_clzll(p);
    const uint64_t mask = (top >= 63) ? ~0ULL : ((1ULL << (top + 1)) - 1);
    W *dst = out + cm * degree;
    for (size_t i = 0; i < degree; ++i) {
      uint64_t v = (uint64_t)stream_value<W>(seed, cm * degree + i) & mask;
      if (v >= p) v -= p;
      dst[i] = (W)v;
    }
  }
  return YA_OK;
}

// poly::set(gaussian) (include/yell/details/poly.hpp:236-260): coefficient i
// takes stream words 2i (magnitude, from the table) and 2i+1 (sign, top
// bit); the signed sample r is shared by all residues, r < 0 ? p + r : r.
// |r| <= ceil(13 sigma) is below every prime of both builds.
// Signed samples 0..count-1 of the stream of `seed`: word 2i against the
// cumulative table (the magnitude), the top bit of word 2i+1 (the sign).
int gaussian_signed(int64_t *r, size_t count, double sigma, uint64_t seed) {
  std::vector<uint64_t> cdf;
  int rc = ya::gaussian_cdt(sigma, cdf);
  if (rc) return rc;
  const size_t tail = cdf.size() - 1;
  for (size_t i = 0; i < count; ++i) {
    const uint64_t u = ya::stream_word(seed, 2 * i), s = ya::stream_word(seed, 2 * i + 1);
    size_t k = 0;
    while (k < tail && u >= cdf[k]) ++k;
    r[i] = (k != 0 && (s >> 63)) ? -(int64_t)k : (int64_t)k;
  }
  return YA_OK;
}

template <class W, class PF>
int sample_gaussian(W *out, size_t degree, size_t nmoduli, double sigma, uint64_t seed, PF prime) {
  if (!out) return ya::fail(YA_ERR_INVALID, "null output");
  std::vector<int64_t> r(degree);
  int rc = gaussian_signed(r.data(), degree, sigma, seed);
  if (rc) return rc;
  for (size_t cm = 0; cm < nmoduli; ++cm) {
    const int64_t p = (int64_t)prime(cm);
    W *dst = out + cm * degree;
    for (size_t i = 0; i < degree; ++i) dst[i] = (W)(r[i] < 0 ? p + r[i] : r[i]);
  }
  return YA_OK;
}

// poly::set(ZO_dist) (include/yell/details/poly.hpp:267-279): byte i of the
// stream is coefficient i's r, shared by all residues;
// r <= rho ? ((r & 2) ? p-1 : 1) : 0.
template <class W, class PF>
int sample_zo(W *out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed, PF prime) {
  if (!out || rho > 255) return ya::fail(YA_ERR_INVALID, "null output or rho > 255");
  std::vector<uint8_t> r(degree);
  for (size_t i = 0; i < degree; ++i) r[i] = (uint8_t)(ya::stream_word(seed, i >> 3) >> (8 * (i & 7)));
  for (size_t cm = 0; cm < nmoduli; ++cm) {
    const W p = (W)prime(cm);
    for (size_t i = 0; i < degree; ++i) out[cm * degree + i] = r[i] <= rho ? ((r[i] & 2) ? p - 1 : 1) : 0;
  }
  return YA_OK;
}

// poly::set(hwt_dist) (include/yell/details/poly.hpp:287-330): slots
// hit[0..hwt) start at positions 0..hwt-1; for k = hwt .. degree-1 a position
// pos uniform below k (ya::reservoir_pos) replaces slot pos if pos < hwt.
// The hit positions are sorted and the j-th smallest gets +1 if stream word j
// has bit 1 set, else p-1 (-1); all residues share the pattern.
template <class W, class PF>
int sample_hwt(W *out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed, PF prime) {
  if (!out || hwt == 0 || hwt > degree) return ya::fail(YA_ERR_INVALID, "hwt must be in [1, degree]");
  std::vector<size_t> hit(hwt);
  for (size_t k = 0; k < hwt; ++k) hit[k] = k;
  for (size_t k = hwt; k < degree; ++k) {
    const uint64_t pos = ya::reservoir_pos(seed, k);
    if (pos < hwt) hit[pos] = k;
  }
  std::sort(hit.begin(), hit.end());
  for (size_t cm = 0; cm < nmoduli; ++cm) {
    const W p = (W)prime(cm);
    W *dst = out + cm * degree;
    std::fill(dst, dst + degree, 0);
    for (size_t j = 0; j < hwt; ++j) dst[hit[j]] = (ya::stream_word(seed, j) & 2) ? 1 : p - 1;
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

int ya_gaussian_cdt(double sigma, uint64_t *out, size_t capacity, size_t *count) {
  std::vector<uint64_t> cdf;
  int rc = ya::gaussian_cdt(sigma, cdf);
  if (rc) return rc;
  if (count) *count = cdf.size();
  if (out) {
    if (capacity < cdf.size()) return ya::fail(YA_ERR_INVALID, "capacity %zu < %zu table entries", capacity, cdf.size());
    std::copy(cdf.begin(), cdf.end(), out);
  }
  return YA_OK;
}

int ya_gaussian_noise(int64_t *out, size_t count, double sigma, uint64_t seed) {
  if (!out && count) return ya::fail(YA_ERR_INVALID, "null output");
  return gaussian_signed(out, count, sigma, seed);
}

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
