// yell/prng/FastGaussianNoise.hpp -- drop-in for the reference's
// FastGaussianNoise<in_class, out_class, _lu_depth>
// (include/yell/prng/FastGaussianNoise.hpp:40-76): the same template
// parameters, constructor and getNoise().
//
// The reference builds MPFR lookup tables for a centred discrete Gaussian of
// parameter sigma within 2^-security of the ideal distribution and reads
// /dev/urandom.  This one draws from the library's seeded table sampler
// (ya_gaussian_noise: density proportional to exp(-x^2 / (2 sigma^2)) on
// |x| <= ceil(13 sigma), about 2^-120 of tail mass left out, the next seed of
// yell::next_seed() per call), so the distribution is the reference's and the
// words are reproducible.  Differences a caller sees: only centre 0 is
// accepted (the reference also takes a double or MPFR centre), `security` and
// `samples` are recorded but do not change the table, and `verbose` prints
// nothing.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "yell/prng/seed.hpp"
#include "yell_amd.h"

namespace yell {

template <class in_class, class out_class, unsigned _lu_depth>
class FastGaussianNoise {
  // the reference checks these at run time and exits
  // (FastGaussianNoise.hpp:213-236)
  static_assert(_lu_depth == 1 || _lu_depth == 2, "FastGaussianNoise: _lu_depth must be 1 or 2");
  static_assert(std::is_same<in_class, uint8_t>::value || std::is_same<in_class, uint16_t>::value,
                "FastGaussianNoise: in_class must be uint8_t or uint16_t");
  static_assert(std::is_integral<out_class>::value, "FastGaussianNoise: out_class must be an integer type");

 public:
  FastGaussianNoise(double sigma, unsigned int security, unsigned int samples, double center_d = 0,
                    bool verbose = false)
      : sigma_(sigma), security_(security), samples_(samples) {
    (void)verbose;
    if (center_d != 0) throw std::invalid_argument("FastGaussianNoise: only centre 0 is supported");
    if (!(sigma > 0)) throw std::invalid_argument("FastGaussianNoise: sigma must be positive");
  }

  // rlen signed samples, each stored as out_class (two's complement), as the
  // reference's getNoise fills the buffer that poly::set reads as
  // params::signed_type (include/yell/details/poly.hpp:243-246).
  void getNoise(out_class *const rand_data2out, uint64_t rlen) {
    std::vector<int64_t> s(rlen);
    const int rc = ya_gaussian_noise(s.data(), (size_t)rlen, sigma_, next_seed());
    if (rc != YA_OK) throw std::runtime_error(std::string("FastGaussianNoise::getNoise: ") + ya_last_error());
    for (uint64_t i = 0; i < rlen; ++i) rand_data2out[i] = (out_class)s[i];
  }

  double sigma() const { return sigma_; }
  unsigned int security() const { return security_; }
  unsigned int samples() const { return samples_; }

 private:
  double sigma_;
  unsigned int security_;
  unsigned int samples_;
};

}  // namespace yell
