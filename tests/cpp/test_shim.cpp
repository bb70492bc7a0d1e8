// test_shim.cpp -- exercises the C++ drop-in headers (include/yell/*.hpp)
// on the GPU the way the reference's own users call yell::poly / ntt / ops
// (benchmark_ntt.cc, example.cpp).  Exit status 0 = every check passed.
// Built twice: 64-bit moduli (test_shim) and -DYELL_USE_32BITS_MODULI
// (test_shim32), as the reference's compile-time switch.
#include <array>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "yell/ntt.hpp"
#include "yell/ops.hpp"
#include "yell/poly.hpp"

static int g_fail = 0;
#define CHECK(c, ...)                                   \
  do {                                                  \
    if (!(c)) {                                         \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);                \
      std::fprintf(stderr, "\n");                       \
      ++g_fail;                                         \
    }                                                   \
  } while (0)

using T = yell::params::value_type;

// The reference's Gaussian generator, as example.cpp:5-9 declares it.
using gauss_t = yell::FastGaussianNoise<uint16_t, T, 2>;
using gauss_struct = yell::gaussian<uint16_t, T, 2>;
static gauss_t g_prng(3.2, 128, 1 << 14);

// Poly-multiplication through the reference's call sequence equals the fused
// entry point and a schoolbook negacyclic convolution on sampled coefficients.
static void test_polymul() {
  constexpr size_t N = 4096;
  const size_t L = 3;
  yell::poly<N> a(L, yell::uniform{}), b(L, gauss_struct(&g_prng));
  yell::poly<N> fa(a), fb(b);
  fa.forward();
  fb.forward();
  yell::poly<N> c = fa * fb;
  c.backward();
  yell::poly<N> fused = yell::multiply(a, b);
  CHECK(c == fused, "sequence vs fused poly-multiplication differ");
  const auto ha = a.to_host(), hb = b.to_host(), hc = c.to_host();
  yell::ops::mulmod mul;
  yell::ops::addmod add;
  yell::ops::submod sub;
  for (size_t cm = 0; cm < L; ++cm) {
    for (size_t k : {size_t(0), size_t(1), size_t(777), N - 1}) {
      T pos = 0, neg = 0;
      for (size_t i = 0; i < N; ++i) {
        const T prod = mul(ha[cm * N + i], hb[cm * N + (i <= k ? k - i : N + k - i)], cm);
        if (i <= k) pos = add(pos, prod, cm);
        else neg = add(neg, prod, cm);
      }
      CHECK(hc[cm * N + k] == sub(pos, neg, cm), "schoolbook mismatch cm=%zu k=%zu", cm, k);
    }
  }
}

// ntt<N> on host memory (staged) equals the poly's device transform; tables
// satisfy the reference's identities.
static void test_ntt_host_pointer_and_tables() {
  constexpr size_t N = 1024;
  yell::poly<N> a(2, yell::uniform{});
  auto h = a.to_host();
  std::vector<T> r0(h.begin(), h.begin() + N);
  yell::ntt<N>::forward_lazy(r0.data(), 0);
  a.forward_lazy();
  const auto d = a.to_host();
  CHECK(std::equal(r0.begin(), r0.end(), d.begin()), "host-staged forward_lazy differs");
  yell::ntt<N>::backward(a.ptr_at(1), 1);  // device pointer path, in place
  const auto t = yell::ntt<N>::init_table(1);
  yell::ops::mulmod mul;
  CHECK(mul(t->invDegree, N, 1) == 1, "invDegree * N != 1");
  CHECK(t->shoupPhiTbl[5] == yell::ops::shoupify(t->phiTbl[5], 1), "Shoup companion");
  CHECK(mul(t->phiTbl[1], t->phiTbl[1], 1) == yell::params::P[1] - 1, "phi^(N/2)^2 != -1");
#ifndef YELL_USE_32BITS_MODULI
  // the tables passed as the reference passes them to its out-of-line seam
  // (details/ntt.hpp:111-118: the arrays decay to pointers)
  const auto t0 = yell::ntt<N>::init_table(0);
  std::vector<T> r1(h.begin(), h.begin() + N);
  CHECK(ya_negacylic_forward_lazy(r1.data(), N, t0->phiTbl, t0->shoupPhiTbl, yell::params::P[0]) == YA_OK,
        "seam call");
  CHECK(r1 == r0, "seam forward_lazy differs from ntt<N>::forward_lazy");
#endif
}

// The toy RLWE scheme of the reference's example.cpp:12-58 at N = 8192,
// L = 4: decryption returns the message up to small noise.
static void test_rlwe_example() {
  constexpr size_t Deg = 8192, NdM = 4;
  yell::poly<Deg> sx(NdM, yell::hwt_dist(64));  // SK
  sx.forward();
  yell::poly<Deg> pbx(NdM), pax(NdM, yell::uniform{});  // PK
  pbx.set(gauss_struct(&g_prng));
  pbx.forward();
  pbx.add_product_of(pax, sx);
  pbx.negate();
  yell::poly<Deg> msg(NdM, yell::uniform{});
  yell::poly<Deg> cbx(NdM), cax(NdM);  // CTXT
  yell::poly<Deg> u(NdM, yell::ZO_dist{});
  u.forward();
  cbx.set(gauss_struct(&g_prng));
  cbx += msg;
  cbx.forward_lazy();
  cbx.add_product_of(u, pbx);
  cax.set(gauss_struct(&g_prng));
  cax.forward_lazy();
  cax.add_product_of(u, pax);
  yell::poly<Deg> plain(NdM);  // decrypt
  plain = cbx;
  plain.add_product_of(sx, cax);
  plain.backward();
  const auto hm = msg.to_host(), hp = plain.to_host();
  long long worst = 0;
  for (size_t cm = 0; cm < NdM; ++cm) {
    const T p = yell::params::P[cm];
    for (size_t i = 0; i < Deg; ++i) {
      const T d = yell::ops::submod{}(hp[cm * Deg + i], hm[cm * Deg + i], cm);
      const long long e = d > p / 2 ? -(long long)(p - d) : (long long)d;
      if (std::llabs(e) > worst) worst = std::llabs(e);
    }
  }
  CHECK(worst < (1ll << 20), "decryption noise %lld too large", worst);
  std::printf("rlwe: max |noise| = %lld\n", worst);

#ifndef YELL_USE_32BITS_MODULI
  // example.cpp:67: the lifted plaintext; each integer reduces to its residues.
  const size_t w = plain.lift_limbs();
  const std::vector<uint64_t> limbs = plain.to_limbs();
  CHECK(limbs.size() == Deg * w, "to_limbs size");
  int bad = 0;
  for (size_t i = 0; i < Deg; i += 97) {
    for (size_t cm = 0; cm < NdM; ++cm) {
      const T p = yell::params::P[cm];
      unsigned __int128 r = 0;  // Horner from the top limb
      for (size_t k = w; k-- > 0;) r = ((r << 64) | limbs[i * w + k]) % p;
      if ((T)r != hp[cm * Deg + i]) ++bad;
    }
  }
  CHECK(bad == 0, "%d lifted coefficients do not reduce to their residues", bad);
#ifdef YELL_AMD_HAVE_GMP
  auto coeffs = plain.poly2mpz();
  for (size_t i = 0; i < Deg; ++i) {
    if (i % 89 == 0) {
      for (size_t cm = 0; cm < NdM; ++cm)
        if (mpz_fdiv_ui(coeffs[i], yell::params::P[cm]) != hp[cm * Deg + i]) ++bad;
    }
    mpz_clear(coeffs[i]);
  }
  CHECK(bad == 0, "poly2mpz: %d coefficients do not reduce to their residues", bad);
  std::printf("rlwe: poly2mpz checked (GMP)\n");
#endif
#endif  // YELL_USE_32BITS_MODULI
}

// poly::set draws on the device; the words equal the host sampler's for the
// seed that set() consumed (next_seed() returns the seed, then advances it).
static void test_device_samplers() {
  constexpr size_t N = 2048, L = 3;
  std::vector<T> h(N * L);
  const uint64_t s0 = 777;
  yell::set_seed(s0);
  const uint64_t first = yell::next_seed();
  const uint64_t step = yell::next_seed() - first;  // the seed increment
  yell::set_seed(s0);
  yell::poly<N> u(L, yell::uniform{}), z(L, yell::ZO_dist{}), w(L, yell::hwt_dist(64));
  yell::poly<N> g(L, gauss_struct(&g_prng));
#ifdef YELL_USE_32BITS_MODULI
  auto host = [&](int k, uint64_t seed) {
    return k == 0 ? ya32_sample_uniform(h.data(), N, L, seed) : k == 1 ? ya32_sample_zo(h.data(), N, L, 0x7F, seed)
           : k == 2 ? ya32_sample_hwt(h.data(), N, L, 64, seed) : ya32_sample_gaussian(h.data(), N, L, 3.2, seed);
  };
#else
  auto host = [&](int k, uint64_t seed) {
    return k == 0 ? ya_sample_uniform(h.data(), N, L, seed) : k == 1 ? ya_sample_zo(h.data(), N, L, 0x7F, seed)
           : k == 2 ? ya_sample_hwt(h.data(), N, L, 64, seed) : ya_sample_gaussian(h.data(), N, L, 3.2, seed);
  };
#endif
  const yell::poly<N> *polys[4] = {&u, &z, &w, &g};
  for (int k = 0; k < 4; ++k) {
    CHECK(host(k, s0 + k * step) == YA_OK, "host sampler %d", k);
    CHECK(polys[k]->to_host() == h, "device sampler %d differs from the host sampler", k);
  }
}

// operator<< prints the reference's text form (details/poly.hpp:336-353).
static void test_stream_output() {
  constexpr size_t N = 8, L = 2;
  yell::poly<N> p(L);
  std::vector<T> h(N * L);
  for (size_t k = 0; k < h.size(); ++k) h[k] = (T)(k % N + 10 * (k / N));
  p.from_host(h);
  std::ostringstream os;
  os << p;
  const std::string want = "{{0ULL, 1ULL, 2ULL, 3ULL, 4ULL, 5ULL, 6ULL, 7ULL}\n"
                           "{10ULL, 11ULL, 12ULL, 13ULL, 14ULL, 15ULL, 16ULL, 17ULL}}";
  CHECK(os.str() == want, "operator<< printed %s", os.str().c_str());
  // recast_as_array views residue cm in place (a device address here)
  const yell::poly<N> &cp = p;
  CHECK((void *)yell::recast_as_array(p, 1) == (void *)p.ptr_at(1), "recast_as_array");
  CHECK((const void *)yell::recast_as_array(cp, 1) == (const void *)cp.cptr_at(1), "recast_as_array const");
  std::array<T, N> back;
  CHECK(ya_memcpy_dtoh(back.data(), yell::recast_as_array(cp, 1), sizeof back, nullptr) == YA_OK, "copy");
  CHECK(back[3] == 13, "recast_as_array residue 1 word 3 = %llu", (unsigned long long)back[3]);
}

// Element access through the proxy of operator()(cm, i), raw_data(), logn,
// the index checks, getNoise and the amplifier of gaussian<In, Out, D>.
static void test_accessors() {
  constexpr size_t N = 1024;
  static_assert(yell::poly<N>::logn == 10, "logn");
  yell::poly<N> p(2);
  p(1, 5) = 123;
  p(0, 0) = p(1, 5);
  p(0, 1) += 7;
  const yell::poly<N> &cp = p;
  CHECK(cp(1, 5) == 123 && (T)p(0, 0) == 123 && cp(0, 1) == 7, "proxy write / read");
  CHECK(p.raw_data()[1] == p.ptr_at(1) && cp.raw_data()[0] == cp.cptr_at(0), "raw_data");
  yell::poly<N> q(std::move(p));
  CHECK(q.raw_data()[1] == q.ptr_at(1) && q(1, 5) == 123, "raw_data after a move");
  int thrown = 0;
  try { (void)q(0, N); } catch (const std::out_of_range &) { ++thrown; }
  try { (void)cp(0, N); } catch (const std::out_of_range &) { ++thrown; }
  try { q(2, 0) = 1; } catch (const std::out_of_range &) { ++thrown; }
  CHECK(thrown == 3, "out-of-range indices: %d of 3 thrown", thrown);

  // getNoise: the signed samples of the stream of the next seed
  yell::set_seed(99);
  std::vector<T> noise(100);
  g_prng.getNoise(noise.data(), noise.size());
  std::vector<int64_t> want(100);
  CHECK(ya_gaussian_noise(want.data(), want.size(), 3.2, 99) == YA_OK, "ya_gaussian_noise");
  bool same = true;
  for (size_t i = 0; i < noise.size(); ++i) same = same && noise[i] == (T)want[i];
  CHECK(same, "getNoise differs from ya_gaussian_noise");

  // gaussian(prng, amplifier): r * amplifier mapped to each residue
  yell::set_seed(5);
  yell::poly<N> g(2, gauss_struct(&g_prng, 3));
  std::vector<int64_t> r(N);
  CHECK(ya_gaussian_noise(r.data(), N, 3.2, 5) == YA_OK, "ya_gaussian_noise");
  const auto h = g.to_host();
  int bad = 0;
  for (size_t cm = 0; cm < 2; ++cm)
    for (size_t i = 0; i < N; ++i) {
      const yell::params::signed_type a = (yell::params::signed_type)(r[i] * 3);
      const T w = a < 0 ? (T)(yell::params::P[cm] + a) : (T)a;
      bad += h[cm * N + i] != w;
    }
  CHECK(bad == 0, "amplified gaussian: %d words differ", bad);
}

#ifndef YELL_USE_32BITS_MODULI
// The reference's smallest and largest rings: N = 4 (its own layer order,
// including its backward there) and N = 2^19, the fused entry point against
// the call sequence.
template <size_t N>
static void check_sequence_vs_fused() {
  const size_t L = 2;
  yell::poly<N> a(L, yell::uniform{}), b(L, gauss_struct(&g_prng));
  yell::poly<N> fa(a), fb(b);
  fa.forward();
  fb.forward();
  yell::poly<N> c = fa * fb;
  c.backward();
  CHECK(c == yell::multiply(a, b), "N=%zu: sequence vs fused poly-multiplication differ", N);
}
#endif

int main() {
  if (ya_init(0) != YA_OK) {
    std::fprintf(stderr, "no device: %s\n", ya_last_error());
    return 2;
  }
  yell::set_seed(42);
  test_device_samplers();
  yell::set_seed(42);
  test_polymul();
  test_ntt_host_pointer_and_tables();
  test_rlwe_example();
  test_stream_output();
  test_accessors();
#ifndef YELL_USE_32BITS_MODULI
  check_sequence_vs_fused<4>();
  check_sequence_vs_fused<(1u << 19)>();
#endif
  std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "ok", g_fail);
  return g_fail ? 1 : 0;
}
