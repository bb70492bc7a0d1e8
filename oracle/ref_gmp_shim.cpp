// ref_gmp_shim.cpp -- extern "C" entry points over the reference's CRT lift.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (by oracle/Makefile) together with the
// reference's lib/gmp.cpp, lib/mem.cpp, lib/ntt.cpp, lib/ops.cpp,
// lib/math.cpp and lib/params.cpp, read in place under the reference
// checkout, into oracle/_ref/libyellref_gmp.so, linked against the system
// GMP.  It fills a reference poly<N> residue by residue (ptr_at), calls its
// poly2mpz() (include/yell/details/poly.hpp:162-165 -> gmp.hpp:41-70) and
// exports each mpz as little-endian 64-bit limbs; and it exports the
// gmp::init_table constants (lib/gmp.cpp:6-51).  Used to pin the GPU lift
// and the Python restatement of it.
#include "yell/poly.hpp"

#include <chrono>
#include <cstring>
#include <type_traits>

using T = yell::params::value_type;

namespace {

// Writes |z| as `limbs` little-endian words; -1 if it needs more.
int export_limbs(const mpz_t z, uint64_t *out, size_t limbs) {
  size_t n = (mpz_sizeinbase(z, 2) + 63) / 64;
  if (mpz_sgn(z) == 0) n = 0;
  if (n > limbs) return -1;
  std::memset(out, 0, limbs * 8);
  size_t count = 0;
  mpz_export(out, &count, -1, 8, 0, 0, z);
  return 0;
}

template <size_t N>
int lift_n(const T *res, size_t L, uint64_t *out, size_t limbs) {
  yell::poly<N> p(L);
  for (size_t j = 0; j < L; ++j) std::memcpy(p.ptr_at(j), res + j * N, sizeof(T) * N);
  auto r = p.poly2mpz();
  int rc = 0;
  for (size_t i = 0; i < N; ++i) {
    if (!rc) rc = export_limbs(r[i], out + i * limbs, limbs);
    mpz_clear(r[i]);
  }
  return rc;
}

}  // namespace

extern "C" {

// poly<2^logn>(L) filled from res ([L][N]) -> poly2mpz -> out ([N][limbs]).
int yrefg_poly2mpz(const T *res, int logn, size_t L, uint64_t *out, size_t limbs) {
  switch (logn) {
    case 3: return lift_n<(1u << 3)>(res, L, out, limbs);
    case 4: return lift_n<(1u << 4)>(res, L, out, limbs);
    case 5: return lift_n<(1u << 5)>(res, L, out, limbs);
    case 6: return lift_n<(1u << 6)>(res, L, out, limbs);
    case 7: return lift_n<(1u << 7)>(res, L, out, limbs);
    case 8: return lift_n<(1u << 8)>(res, L, out, limbs);
    case 9: return lift_n<(1u << 9)>(res, L, out, limbs);
    case 10: return lift_n<(1u << 10)>(res, L, out, limbs);
    case 11: return lift_n<(1u << 11)>(res, L, out, limbs);
    case 12: return lift_n<(1u << 12)>(res, L, out, limbs);
    case 13: return lift_n<(1u << 13)>(res, L, out, limbs);
    case 14: return lift_n<(1u << 14)>(res, L, out, limbs);
    case 15: return lift_n<(1u << 15)>(res, L, out, limbs);
    case 16: return lift_n<(1u << 16)>(res, L, out, limbs);
    default: return -2;
  }
}

// gmp::init_table(L): moduli_product_, modulus_shoup (2 limbs),
// lifting_integers ([L][limbs]), shift and bits_in_moduli_product.
int yrefg_constants(size_t L, size_t limbs, uint64_t *q, uint64_t *m2, uint64_t *lift, size_t *shift,
                    size_t *bits) {
  auto tbl = yell::gmp::init_table(L);
  if (export_limbs(tbl->moduli_product_, q, limbs)) return -1;
  if (export_limbs(tbl->modulus_shoup, m2, 2)) return -1;
  for (size_t j = 0; j < L; ++j)
    if (export_limbs(tbl->lifting_integers[j], lift + j * limbs, limbs)) return -1;
  *shift = tbl->shift_modulus_shoup;
  *bits = tbl->bits_in_moduli_product;
  return 0;
}

// Wall-clock of `reps` poly2mpz calls on one poly<2^logn>(L) (one thread),
// seconds per call; the residues come from res.
double yrefg_bench_poly2mpz(const T *res, int logn, size_t L, int reps) {
  if (logn != 14 && logn != 16) return -1.0;
  auto run = [&](auto tag) {
    constexpr size_t N = decltype(tag)::value;
    yell::poly<N> p(L);
    for (size_t j = 0; j < L; ++j) std::memcpy(p.ptr_at(j), res + j * N, sizeof(T) * N);
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < reps; ++k) {
      auto r = p.poly2mpz();
      for (size_t i = 0; i < N; ++i) mpz_clear(r[i]);
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count() / reps;
  };
  if (logn == 14) return run(std::integral_constant<size_t, (1u << 14)>());
  return run(std::integral_constant<size_t, (1u << 16)>());
}

}  // extern "C"
