// yell/poly.hpp -- drop-in for the reference's yell::poly<N>
// (include/yell/poly.hpp:43-168, include/yell/details/poly.hpp) whose
// residues live in GPU memory.
//
// Same constructors, samplers, NTT methods, operators and accessors as the
// reference; the differences a caller sees:
//   * storage is one residue-major device buffer [L][N]: ptr_at(cm) and the
//     residue pointers of raw_data() are device pointers (data() is the
//     base), so host code must not dereference them;
//   * element access operator()(cm, i) reads one word across PCIe; on a
//     non-const poly it returns a reference-like proxy, so p(cm, i) = v also
//     writes one word; to_host() / from_host() move whole polynomials;
//   * samplers are seeded (yell::set_seed); yell::gaussian<In, Out, D> wraps
//     a yell::FastGaussianNoise<In, Out, D> as in the reference, drawn on
//     the device from that sampler's sigma (see yell/prng/FastGaussianNoise.hpp);
//   * with -DYELL_USE_32BITS_MODULI the 32-bit build (uint32 words, 30-bit
//     primes, degrees 2^3 .. 2^14) is used, as in the reference;
//   * poly2mpz (the CRT lift, include/yell/gmp.hpp:41-70) runs on the GPU
//     (64-bit build);
//     it is declared when <gmp.h> is on the include path (link -lgmp), and
//     to_limbs() gives the same integers as little-endian 64-bit limbs
//     without GMP;
//   * failures throw std::runtime_error, out-of-range indices
//     std::out_of_range (the reference asserts).
// Every arithmetic result is bit-identical to the reference's.
#pragma once
#include <array>
#include <atomic>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "yell/ntt.hpp"
#include "yell/ops.hpp"
#include "yell/params.hpp"
#include "yell/prng/FastGaussianNoise.hpp"
#include "yell/prng/seed.hpp"
#include "yell_amd.h"
#if __has_include(<gmp.h>)
#include <gmp.h>
#define YELL_AMD_HAVE_GMP 1
#endif

namespace yell {

namespace detail {  // poly entry points of the build params::value_type selects
#define YELL_AMD_API2(NAME, F64, F32)                                                                   \
  inline int NAME(uint64_t *x, size_t n, size_t l) { return F64(x, n, l, nullptr); }                 \
  inline int NAME(uint32_t *x, size_t n, size_t l) { return F32(x, n, l, nullptr); }
YELL_AMD_API2(api_poly_forward, ya_poly_forward, ya32_poly_forward)
YELL_AMD_API2(api_poly_forward_lazy, ya_poly_forward_lazy, ya32_poly_forward_lazy)
YELL_AMD_API2(api_poly_backward, ya_poly_backward, ya32_poly_backward)
YELL_AMD_API2(api_poly_negate, ya_poly_negate, ya32_poly_negate)
#undef YELL_AMD_API2
#define YELL_AMD_API3(NAME, F64, F32)                                                                   \
  inline int NAME(uint64_t *o, const uint64_t *a, const uint64_t *b, size_t n, size_t l, void *s) {  \
    return F64(o, a, b, n, l, s);                                                                     \
  }                                                                                                   \
  inline int NAME(uint32_t *o, const uint32_t *a, const uint32_t *b, size_t n, size_t l, void *s) {  \
    return F32(o, a, b, n, l, s);                                                                     \
  }
YELL_AMD_API3(api_poly_mul, ya_poly_mul, ya32_poly_mul)
YELL_AMD_API3(api_poly_add, ya_poly_add, ya32_poly_add)
YELL_AMD_API3(api_poly_sub, ya_poly_sub, ya32_poly_sub)
YELL_AMD_API3(api_poly_add_product_of, ya_poly_add_product_of, ya32_poly_add_product_of)
YELL_AMD_API3(api_poly_multiply, ya_poly_multiply, ya32_poly_multiply)
#undef YELL_AMD_API3
// Device samplers (same seeded streams as the host samplers ya_sample_*).
inline int api_sample_uniform(uint64_t *d, size_t n, size_t l, uint64_t s) {
  return ya_sample_uniform_device(d, n, l, s, nullptr);
}
inline int api_sample_uniform(uint32_t *d, size_t n, size_t l, uint64_t s) {
  return ya32_sample_uniform_device(d, n, l, s, nullptr);
}
inline int api_sample_zo(uint64_t *d, size_t n, size_t l, uint32_t r, uint64_t s) {
  return ya_sample_zo_device(d, n, l, r, s, nullptr);
}
inline int api_sample_zo(uint32_t *d, size_t n, size_t l, uint32_t r, uint64_t s) {
  return ya32_sample_zo_device(d, n, l, r, s, nullptr);
}
inline int api_sample_hwt(uint64_t *d, size_t n, size_t l, uint32_t w, uint64_t s) {
  return ya_sample_hwt_device(d, n, l, w, s, nullptr);
}
inline int api_sample_hwt(uint32_t *d, size_t n, size_t l, uint32_t w, uint64_t s) {
  return ya32_sample_hwt_device(d, n, l, w, s, nullptr);
}
inline int api_sample_gaussian(uint64_t *d, size_t n, size_t l, double g, uint64_t a, uint64_t s) {
  return ya_sample_gaussian_device(d, n, l, g, a, s, nullptr);
}
inline int api_sample_gaussian(uint32_t *d, size_t n, size_t l, double g, uint64_t a, uint64_t s) {
  return ya32_sample_gaussian_device(d, n, l, g, a, s, nullptr);
}
}  // namespace detail

struct uniform {};
struct hwt_dist {
  uint32_t hwt;
  explicit hwt_dist(uint32_t h) : hwt(h) {}
};
struct ZO_dist {
  uint8_t rho;
  explicit ZO_dist(uint8_t r = 0x7F) : rho(r) {}
};
// include/yell/poly.hpp:32-41: the Gaussian generator, a sampler plus the
// amplifier the signed samples are multiplied by.
template <class in_class, class out_class, unsigned _lu_depth>
struct gaussian {
  FastGaussianNoise<in_class, out_class, _lu_depth> *fg_prng;
  uint64_t amplifier;
  explicit gaussian(FastGaussianNoise<in_class, out_class, _lu_depth> *prng) : fg_prng{prng}, amplifier{1} {}
  gaussian(FastGaussianNoise<in_class, out_class, _lu_depth> *prng, uint64_t amp) : fg_prng{prng}, amplifier{amp} {}
};

namespace detail {
constexpr size_t log2_of(size_t n) { return n <= 1 ? 0 : 1 + log2_of(n >> 1); }
// Smallest degree of this build: the 64-bit engine runs the reference's
// N = 4 (its forward needs N >= 4, lib/ntt.cpp:123-135), the 32-bit one 2^3.
#ifdef YELL_USE_32BITS_MODULI
constexpr size_t kMinDegree = 8;
#else
constexpr size_t kMinDegree = 4;
#endif
}  // namespace detail

template <size_t degree_>
class poly {
  // the reference asserts degree_ < kMaxPolyDegree (include/yell/poly.hpp:45)
  static_assert(degree_ >= detail::kMinDegree && degree_ < params::kMaxPolyDegree && (degree_ & (degree_ - 1)) == 0,
                "degree must be a power of two of at least 4 (8 with 32-bit moduli) below params::kMaxPolyDegree");

 public:
  using value_type = params::value_type;
  using gt_value_type = params::gt_value_type;
  using signed_type = params::signed_type;
  using pointer_type = value_type *;
  using const_pointer_type = const value_type *;
  using iterator = pointer_type;
  using const_iterator = const_pointer_type;
  static constexpr size_t degree = degree_;
  static constexpr size_t logn = detail::log2_of(degree_);

  explicit poly(size_t nmoduli) : nmoduli_(0), d_(nullptr) { resize_moduli_count(nmoduli); clear(); }
  explicit poly(size_t nmoduli, const uniform &m) : poly(nmoduli) { set(m); }
  explicit poly(size_t nmoduli, const hwt_dist &m) : poly(nmoduli) { set(m); }
  explicit poly(size_t nmoduli, const ZO_dist &m) : poly(nmoduli) { set(m); }
  template <class in_class, unsigned _lu_depth>
  explicit poly(size_t nmoduli, const gaussian<in_class, value_type, _lu_depth> &m) : poly(nmoduli) {
    set(m);
  }
  poly(const poly &o) : poly(o.nmoduli_) { copy_from(o); }
  poly(poly &&o) noexcept : nmoduli_(o.nmoduli_), d_(o.d_), ptrs_(std::move(o.ptrs_)) {
    o.d_ = nullptr;
    o.nmoduli_ = 0;
    o.ptrs_.clear();
  }
  poly &operator=(const poly &o) {
    if (this != &o) {
      resize_moduli_count(o.nmoduli_);
      copy_from(o);
    }
    return *this;
  }
  ~poly() { if (d_) ya_free(d_); }

  friend void swap(poly &a, poly &b) {
    std::swap(a.d_, b.d_);
    std::swap(a.nmoduli_, b.nmoduli_);
    std::swap(a.ptrs_, b.ptrs_);
  }

  // ---- samplers (details/poly.hpp:200-330), drawn on the device into the
  // residues (seeded; the host samplers ya_sample_* give the same words).
  void set(const uniform &) {
    sample([&](uint64_t s) { return detail::api_sample_uniform(d_, degree, nmoduli_, s); });
  }
  void set(const ZO_dist &m) {
    sample([&](uint64_t s) { return detail::api_sample_zo(d_, degree, nmoduli_, m.rho, s); });
  }
  void set(const hwt_dist &m) {
    sample([&](uint64_t s) { return detail::api_sample_hwt(d_, degree, nmoduli_, m.hwt, s); });
  }
  // details/poly.hpp:236-260: one signed sample per coefficient from the
  // generator's distribution, r *= amplifier in params::signed_type, then
  // r < 0 ? p + r : r in every residue; drawn on the device with the
  // sampler's sigma and the next seed.
  template <class in_class, unsigned _lu_depth>
  void set(const gaussian<in_class, value_type, _lu_depth> &m) {
    if (!m.fg_prng) throw std::invalid_argument("gaussian: null FastGaussianNoise");
    const double sigma = m.fg_prng->sigma();
    sample([&](uint64_t s) { return detail::api_sample_gaussian(d_, degree, nmoduli_, sigma, m.amplifier, s); });
  }

  // ---- access
  iterator ptr_at(size_t cm) { check_cm(cm); return d_ + cm * degree; }
  iterator ptr_end(size_t cm) { return ptr_at(cm) + degree; }
  const_iterator cptr_at(size_t cm) const { check_cm(cm); return d_ + cm * degree; }
  const_iterator cptr_end(size_t cm) const { return cptr_at(cm) + degree; }
  value_type *data() { return d_; }
  const value_type *data() const { return d_; }
  // include/yell/poly.hpp:119-121: the residue pointers (device addresses).
  value_type **raw_data() { return ptrs_.data(); }
  value_type *const *raw_data() const { return ptrs_.data(); }
  size_t moduli_count() const { return nmoduli_; }
  value_type get_modulus(size_t cm) const { check_cm(cm); return params::P[cm]; }

  // One word of residue cm, as the reference's value_type& of
  // include/yell/poly.hpp:105-108: reads on conversion, writes on assignment.
  class word_ref {
   public:
    word_ref(value_type *dev) : dev_(dev) {}
    operator value_type() const {
      value_type v;
      ya_check(ya_memcpy_dtoh(&v, dev_, sizeof v, nullptr), "read word");
      return v;
    }
    word_ref &operator=(value_type v) {
      ya_check(ya_memcpy_htod(dev_, &v, sizeof v, nullptr), "write word");
      return *this;
    }
    word_ref &operator=(const word_ref &o) { return *this = static_cast<value_type>(o); }
    word_ref &operator+=(value_type v) { return *this = static_cast<value_type>(*this) + v; }
    word_ref &operator-=(value_type v) { return *this = static_cast<value_type>(*this) - v; }
    friend std::ostream &operator<<(std::ostream &os, const word_ref &r) { return os << static_cast<value_type>(r); }

   private:
    value_type *dev_;
  };
  word_ref operator()(size_t cm, size_t i) { return word_ref(ptr_at(cm) + check_i(i)); }
  value_type operator()(size_t cm, size_t i) const {
    value_type v;
    ya_check(ya_memcpy_dtoh(&v, cptr_at(cm) + check_i(i), sizeof v, nullptr), "read word");
    return v;
  }
  void set_word(size_t cm, size_t i, value_type v) { (*this)(cm, i) = v; }
  std::vector<value_type> to_host() const {
    std::vector<value_type> h(degree * nmoduli_);
    if (!h.empty()) ya_check(ya_memcpy_dtoh(h.data(), d_, h.size() * sizeof(value_type), nullptr), "to_host");
    return h;
  }
  void from_host(const std::vector<value_type> &h) {
    if (h.size() != degree * nmoduli_) throw std::runtime_error("from_host: size mismatch");
    if (!h.empty()) ya_check(ya_memcpy_htod(d_, h.data(), h.size() * sizeof(value_type), nullptr), "from_host");
  }

#ifndef YELL_USE_32BITS_MODULI
  // ---- CRT lift (details/poly.hpp:162-165): coefficient i of the result is
  // words [i * lift_limbs(), (i + 1) * lift_limbs()), little-endian.
  size_t lift_limbs() const {
    size_t w = 0;
    ya_check(ya_crt_limbs(nmoduli_, &w), "ya_crt_limbs");
    return w;
  }
  std::vector<uint64_t> to_limbs() const {
    const size_t w = lift_limbs();
    std::vector<uint64_t> h(degree * w);
    uint64_t *d = nullptr;
    ya_check(ya_malloc((void **)&d, h.size() * 8), "ya_malloc");
    int rc = ya_poly_crt_lift(d, d_, degree, nmoduli_, nullptr);
    if (rc == YA_OK) rc = ya_memcpy_dtoh(h.data(), d, h.size() * 8, nullptr);
    ya_free(d);
    ya_check(rc, "ya_poly_crt_lift");
    return h;
  }
#ifdef YELL_AMD_HAVE_GMP
  // The caller owns the returned integers (mpz_clear each), as in the reference.
  std::array<mpz_t, degree_> poly2mpz() const {
    const size_t w = lift_limbs();
    const std::vector<uint64_t> h = to_limbs();
    std::array<mpz_t, degree_> rop;
    for (size_t i = 0; i < degree_; ++i) {
      mpz_init2(rop[i], 64 * w);
      mpz_import(rop[i], w, -1, sizeof(uint64_t), 0, 0, h.data() + i * w);
    }
    return rop;
  }
#endif
#endif  // YELL_USE_32BITS_MODULI

  bool operator==(const poly &o) const { return nmoduli_ == o.nmoduli_ && to_host() == o.to_host(); }
  bool operator!=(const poly &o) const { return !(*this == o); }

  // ---- NTT (details/poly.hpp:170-186)
  void forward() { if (nmoduli_) ya_check(detail::api_poly_forward(d_, degree, nmoduli_), "forward"); }
  void forward_lazy() { if (nmoduli_) ya_check(detail::api_poly_forward_lazy(d_, degree, nmoduli_), "forward_lazy"); }
  void backward() { if (nmoduli_) ya_check(detail::api_poly_backward(d_, degree, nmoduli_), "backward"); }

  // ---- arithmetic (details/poly.hpp:118-158)
  poly &add_product_of(const poly &a, const poly &b) {
    same(a);
    same(b);
    ya_check(detail::api_poly_add_product_of(d_, a.d_, b.d_, degree, nmoduli_, nullptr), "add_product_of");
    return *this;
  }
  void negate() { if (nmoduli_) ya_check(detail::api_poly_negate(d_, degree, nmoduli_), "negate"); }
  // a * s computed as a * r + a * (s - r) with a fresh uniform r; the masks
  // are cleared before their memory is released (details/poly.hpp:134-150).
  poly &mul_sensitive_poly(const poly &s) {
    same(s);
    poly r0(nmoduli_, uniform{});
    poly r1(s);
    r1 -= r0;
    r1 *= *this;
    *this *= r0;
    *this += r1;
    r0.clear();
    r1.clear();
    return *this;
  }
  void clear() { if (nmoduli_) ya_check(ya_memset_zero(d_, bytes(), nullptr), "clear"); }
  void resize_moduli_count(size_t n) {
    if (n > params::kMaxNbModuli) throw std::runtime_error("too many moduli");
    if (n == nmoduli_) return;
    value_type *nd = nullptr;
    if (n) {
      void *p = nullptr;
      if (ya_malloc(&p, n * degree * sizeof(value_type)) != YA_OK)
        throw std::runtime_error("Memory allocation failed.");
      nd = static_cast<value_type *>(p);
      const size_t keep = std::min(n, nmoduli_);
      if (keep) ya_check(ya_memcpy_dtod(nd, d_, keep * degree * sizeof(value_type), nullptr), "resize copy");
      if (n > keep) ya_check(ya_memset_zero(nd + keep * degree, (n - keep) * degree * sizeof(value_type), nullptr), "resize zero");
    }
    if (d_) ya_free(d_);
    d_ = nd;
    nmoduli_ = n;
    ptrs_.resize(n);
    for (size_t cm = 0; cm < n; ++cm) ptrs_[cm] = d_ + cm * degree;
  }
  void check_vaild_range() const {
    const auto h = to_host();
    for (size_t cm = 0; cm < nmoduli_; ++cm)
      for (size_t i = 0; i < degree; ++i)
        if (h[cm * degree + i] >= params::P[cm]) throw std::runtime_error("coefficient out of range");
  }

 private:
  template <class F>
  void sample(F f) {
    if (nmoduli_) ya_check(f(next_seed()), "sampler");
  }
  void copy_from(const poly &o) {
    if (nmoduli_) ya_check(ya_memcpy_dtod(d_, o.d_, bytes(), nullptr), "copy");
  }
  size_t bytes() const { return nmoduli_ * degree * sizeof(value_type); }
  void check_cm(size_t cm) const { if (cm >= nmoduli_) throw std::out_of_range("modulus index"); }
  static size_t check_i(size_t i) {
    if (i >= degree) throw std::out_of_range("coefficient index");
    return i;
  }
  void same(const poly &o) const {
    if (o.nmoduli_ != nmoduli_) throw std::runtime_error("moduli counts differ");
  }

  size_t nmoduli_;
  value_type *d_;
  std::vector<value_type *> ptrs_;  // residue cm at d_ + cm * degree (raw_data)
};

// Element-wise operators (include/yell/ops.hpp:7-41, poly.hpp:163-168): the
// binary forms allocate the result, the compound forms work in place; both
// use min(moduli counts) residues as the reference does.
#define YELL_AMD_BINARY(OP, FN)                                                       \
  template <size_t N>                                                                 \
  poly<N> OP(const poly<N> &a, const poly<N> &b) {                                    \
    const size_t L = std::min(a.moduli_count(), b.moduli_count());                    \
    poly<N> r(L);                                                                     \
    if (L) ya_check(FN(r.data(), a.data(), b.data(), N, L, nullptr), #FN);            \
    return r;                                                                         \
  }
#define YELL_AMD_SELF(OP, FN)                                                         \
  template <size_t N>                                                                 \
  poly<N> &OP(poly<N> &a, const poly<N> &b) {                                         \
    const size_t L = std::min(a.moduli_count(), b.moduli_count());                    \
    if (L) ya_check(FN(a.data(), a.data(), b.data(), N, L, nullptr), #FN);            \
    return a;                                                                         \
  }
YELL_AMD_BINARY(operator*, detail::api_poly_mul)
YELL_AMD_BINARY(operator+, detail::api_poly_add)
YELL_AMD_BINARY(operator-, detail::api_poly_sub)
YELL_AMD_SELF(operator*=, detail::api_poly_mul)
YELL_AMD_SELF(operator+=, detail::api_poly_add)
YELL_AMD_SELF(operator-=, detail::api_poly_sub)
#undef YELL_AMD_BINARY
#undef YELL_AMD_SELF

// No counterpart in the reference: the whole poly-multiplication
// backward(forward(a) * forward(b)) in one fused launch, a and b untouched.
template <size_t N>
poly<N> multiply(const poly<N> &a, const poly<N> &b) {
  const size_t L = std::min(a.moduli_count(), b.moduli_count());
  poly<N> r(L);
  if (L) ya_check(detail::api_poly_multiply(r.data(), a.data(), b.data(), N, L, nullptr), "poly multiply");
  return r;
}

// include/yell/poly.hpp:144-161: residue cm viewed as one std::array.  Here,
// as ptr_at(cm), the pointer is a device address: for device code, or for
// copies through the C ABI.
template <typename value_type, size_t degree>
struct ArrayPointer {
  using T = std::array<value_type, degree> *;
  using cT = const T;
};

template <size_t degree>
typename ArrayPointer<params::value_type, degree>::T recast_as_array(poly<degree> &op, size_t cm) {
  if (cm >= op.moduli_count()) throw std::out_of_range("modulus index");
  return reinterpret_cast<typename ArrayPointer<params::value_type, degree>::T>(op.ptr_at(cm));
}

template <size_t degree>
typename ArrayPointer<params::value_type, degree>::cT recast_as_array(const poly<degree> &op, size_t cm) {
  if (cm >= op.moduli_count()) throw std::out_of_range("modulus index");
  return reinterpret_cast<typename ArrayPointer<params::value_type, degree>::cT>(
      const_cast<params::value_type *>(op.cptr_at(cm)));
}

// include/yell/poly.hpp:142 (text format of details/poly.hpp:336-353): one
// brace-enclosed row of "<word>ULL" per residue, rows separated by newlines,
// all inside one more pair of braces.  One copy of the polynomial to the host.
template <size_t N>
std::ostream &operator<<(std::ostream &os, const poly<N> &p) {
  const size_t L = p.moduli_count();
  const auto h = p.to_host();
  os << "{";
  for (size_t cm = 0; cm < L; ++cm) {
    os << "{";
    for (size_t i = 0; i < N; ++i) os << h[cm * N + i] << "ULL" << (i + 1 < N ? ", " : "}");
    if (cm + 1 < L) os << "\n";
  }
  return os << "}";
}

}  // namespace yell
