/*
 * schoolbook_gmp.c -- GMP big-integer negacyclic convolution (spot oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Computes c = a * b in Z[X]/(X^N + 1) over the
 * integers with mpz arithmetic (no modular reduction inside the sums), then
 * reduces each coefficient into [0, p).  Independent of the 128-bit
 * arithmetic used by yell_oracle.c; built only where gmp.h is present.
 */
#include <gmp.h>
#include <stddef.h>
#include <stdint.h>

static void set_u64(mpz_t z, uint64_t v)
{
  mpz_set_ui(z, (unsigned long)(v >> 32));
  mpz_mul_2exp(z, z, 32);
  mpz_add_ui(z, z, (unsigned long)(v & 0xffffffffu));
}

int or_gmp_schoolbook(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree,
                      uint64_t p)
{
  mpz_t acc, x, y, mp;
  mpz_inits(acc, x, y, mp, NULL);
  set_u64(mp, p);
  for (size_t k = 0; k < degree; ++k) {
    mpz_set_ui(acc, 0);
    for (size_t i = 0; i < degree; ++i) {
      set_u64(x, a[i]);
      if (i <= k) {
        set_u64(y, b[k - i]);
        mpz_addmul(acc, x, y);
      } else {
        set_u64(y, b[degree + k - i]);
        mpz_submul(acc, x, y);
      }
    }
    mpz_fdiv_r(acc, acc, mp);
    uint64_t hi = 0, lo = 0;
    mpz_fdiv_q_2exp(x, acc, 32);
    hi = (uint64_t)mpz_get_ui(x);
    mpz_fdiv_r_2exp(y, acc, 32);
    lo = (uint64_t)mpz_get_ui(y);
    c[k] = (hi << 32) | lo;
  }
  mpz_clears(acc, x, y, mp, NULL);
  return 0;
}
