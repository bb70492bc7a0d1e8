/*
 * yell_oracle.c -- CPU restatement of the reference's NTT / modular arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke entry and
 * the CPU-baseline leg of bench.py load it.  See yell_oracle.h for the list of
 * properties each function is pinned by.
 */
#include "yell_oracle.h"

#include <stdlib.h>
#include <string.h>

#include "../yell_amd/csrc/params_table.inc"

static const uint64_t kP[YA_NUM_MODULI] = YA_TABLE_P;
static const uint64_t kPn[YA_NUM_MODULI] = YA_TABLE_PN;
static const uint64_t kRoot[YA_NUM_MODULI] = YA_TABLE_ROOT;

/* params_64.hpp:21 kMaxPolyDegree = 2^20 */
#define OR_LOG_MAX_DEGREE 20

size_t or_num_moduli(void) { return YA_NUM_MODULI; }
uint64_t or_prime(size_t cm) { return kP[cm]; }
uint64_t or_barrett_const(size_t cm) { return kPn[cm]; }
uint64_t or_root(size_t cm) { return kRoot[cm]; }

/* lib/ops.cpp:5-21.  q = Pn * (x >> 64) + (x << 2)  (mod 2^128);
 * r = lo64(x) - hi64(q) * p (mod 2^64); one conditional subtraction. */
uint64_t or_barrett(or_u128 x, size_t cm)
{
  const or_u128 pn = kPn[cm];
  const or_u128 q = pn * (x >> 64) + (x << 2);
  const uint64_t p = kP[cm];
  uint64_t r = (uint64_t)x - (uint64_t)(q >> 64) * p;
  if (r >= p) r -= p;
  return r;
}

uint64_t or_barrett_hl(uint64_t hi, uint64_t lo, size_t cm)
{
  return or_barrett(((or_u128)hi << 64) | lo, cm);
}

/* lib/ops.cpp:23-27: floor(x * 2^64 / p). */
uint64_t or_shoupify(uint64_t x, size_t cm)
{
  return (uint64_t)(((or_u128)x << 64) / kP[cm]);
}

/* include/yell/ops.hpp:103-120 */
uint64_t or_mulmod(uint64_t x, uint64_t y, size_t cm)
{
  return or_barrett((or_u128)x * y, cm);
}

/* include/yell/ops.hpp:122-142: q = hi64(x*y') * p; r = x*y - q (mod 2^64);
 * mod_correct (ops.hpp:46-52). */
uint64_t or_mulmod_shoup(uint64_t x, uint64_t y, uint64_t yprime, size_t cm)
{
  const uint64_t p = kP[cm];
  const uint64_t q = (uint64_t)(((or_u128)x * yprime) >> 64) * p;
  uint64_t r = x * y - q;
  if (r >= p) r -= p;
  return r;
}

/* include/yell/ops.hpp:54-68 */
uint64_t or_addmod(uint64_t x, uint64_t y, size_t cm)
{
  const uint64_t p = kP[cm];
  uint64_t r = x + y;
  if (r >= p) r -= p;
  return r;
}

/* include/yell/ops.hpp:70-83 */
uint64_t or_submod(uint64_t x, uint64_t y, size_t cm)
{
  const uint64_t p = kP[cm];
  uint64_t r = x + (p - y);
  if (r >= p) r -= p;
  return r;
}

/* include/yell/ops.hpp:144-158: (x*y + rop) over 128 bits, then Barrett. */
uint64_t or_muladd(uint64_t rop, uint64_t x, uint64_t y, size_t cm)
{
  or_u128 v = (or_u128)x * y;
  v += rop;
  return or_barrett(v, cm);
}

/* lib/ntt.cpp:69-83 (Cooley-Tukey, inputs/outputs in [0, 4p)). */
void or_ct_butterfly(uint64_t *x0, uint64_t *x1, uint64_t w, uint64_t wprime, uint64_t p)
{
  const uint64_t two_p = 2 * p;
  uint64_t u0 = *x0;
  const uint64_t u1 = *x1;
  if (u0 >= two_p) u0 -= two_p;
  const uint64_t q = (uint64_t)(((or_u128)u1 * wprime) >> 64);
  const uint64_t t = u1 * w - q * p;
  *x0 = u0 + t;
  *x1 = u0 - t + two_p;
}

/* lib/ntt.cpp:47-63 (Gentleman-Sande, inputs/outputs in [0, 2p)). */
void or_gs_butterfly(uint64_t *x0, uint64_t *x1, uint64_t w, uint64_t wprime, uint64_t p)
{
  const uint64_t two_p = 2 * p;
  const uint64_t u0 = *x0;
  const uint64_t u1 = *x1;
  uint64_t t0 = u0 + u1;
  if (t0 >= two_p) t0 -= two_p;
  const uint64_t t1 = u0 - u1 + two_p;
  const uint64_t q = (uint64_t)(((or_u128)t1 * wprime) >> 64);
  *x0 = t0;
  *x1 = t1 * w - q * p;
}

/* lib/ntt.cpp:89-149.  Layers m = 1, 2, ..., degree/2 with h = degree/(2m);
 * group r of layer m uses wtab[m + r].  The reference's three code blocks
 * (h >= 4 unrolled, h == 2, h == 1) apply the same butterfly to the same
 * pairs; the pairs of one layer are disjoint, so their order is immaterial. */
int or_negacyclic_forward_lazy(uint64_t *x, size_t degree, const uint64_t *wtab,
                               const uint64_t *wtab_shoup, uint64_t p)
{
  if (degree < 4 || (degree & (degree - 1))) return -1;
  size_t m = 1, h = degree >> 1;
  for (; h > 2; m <<= 1, h >>= 1) {               /* :102-120 */
    for (size_t r = 0; r < m; ++r) {
      uint64_t *x0 = x + 2 * h * r, *x1 = x0 + h;
      for (size_t i = 0; i < h; ++i)
        or_ct_butterfly(x0 + i, x1 + i, wtab[m + r], wtab_shoup[m + r], p);
    }
  }
  for (size_t r = 0; r < m; ++r) {                 /* :123-135, h = 2 */
    uint64_t *x0 = x + 4 * r;
    or_ct_butterfly(x0, x0 + 2, wtab[m + r], wtab_shoup[m + r], p);
    or_ct_butterfly(x0 + 1, x0 + 3, wtab[m + r], wtab_shoup[m + r], p);
  }
  m <<= 1;
  for (size_t r = 0; r < m; ++r)                   /* :137-147, h = 1 */
    or_ct_butterfly(x + 2 * r, x + 2 * r + 1, wtab[m + r], wtab_shoup[m + r], p);
  return 0;
}

/* lib/ntt.cpp:151-210.  Layers m = degree/2 (h = 1), degree/4 (h = 2), then
 * m > 1 with h doubling; the m = 1 layer is left to the caller
 * (details/ntt.hpp:151-159), which merges it with the N^-1 scaling. */
int or_negacyclic_backward_lazy(uint64_t *x, size_t degree, const uint64_t *wtab,
                                const uint64_t *wtab_shoup, uint64_t p)
{
  if (degree < 4 || (degree & (degree - 1))) return -1;
  size_t m = degree >> 1, h = 1;
  for (size_t r = 0; r < m; ++r)                   /* :162-172, h = 1 */
    or_gs_butterfly(x + 2 * r, x + 2 * r + 1, wtab[m + r], wtab_shoup[m + r], p);
  m >>= 1;                                         /* :174-187, h = 2 */
  h <<= 1;
  for (size_t r = 0; r < m; ++r) {
    uint64_t *x0 = x + 4 * r;
    or_gs_butterfly(x0, x0 + 2, wtab[m + r], wtab_shoup[m + r], p);
    or_gs_butterfly(x0 + 1, x0 + 3, wtab[m + r], wtab_shoup[m + r], p);
  }
  m >>= 1;                                         /* :189-208 */
  h <<= 1;
  for (; m > 1; m >>= 1, h <<= 1) {
    for (size_t r = 0; r < m; ++r) {
      uint64_t *x0 = x + 2 * h * r, *x1 = x0 + h;
      for (size_t i = 0; i < h; ++i)
        or_gs_butterfly(x0 + i, x1 + i, wtab[m + r], wtab_shoup[m + r], p);
    }
  }
  return 0;
}

/* include/yell/utils/math.hpp:29-43: in-place bit-reversal permutation. */
static void revbin_permute(uint64_t *a, size_t n)
{
  if (n <= 2) return;
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j >= bit; bit >>= 1) j -= bit;
    j += bit;
    if (i < j) {
      uint64_t t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
}

/* include/yell/utils/math.hpp:8-21: a^(p-2) mod p with Barrett mulmod. */
static uint64_t inv_mod_prime(uint64_t a, size_t cm)
{
  uint64_t ret = 1, y = kP[cm] - 2;
  while (y > 0) {
    if (y & 1) ret = or_mulmod(ret, a, cm);
    y >>= 1;
    a = or_mulmod(a, a, cm);
  }
  return ret;
}

static int log2_exact(size_t n)
{
  int l = 0;
  if (n == 0 || (n & (n - 1))) return -1;
  while ((1ull << l) < n) ++l;
  return l;
}

/* include/yell/details/ntt.hpp:27-77 */
int or_tables_init(size_t degree, size_t cm, uint64_t *phi_tbl, uint64_t *phi_shoup,
                   uint64_t *invphi_tbl, uint64_t *invphi_shoup, uint64_t *scalars)
{
  const int logn = log2_exact(degree);
  if (logn < 1 || logn > OR_LOG_MAX_DEGREE || cm >= YA_NUM_MODULI) return -1;
  uint64_t phi = kRoot[cm];
  for (int i = 0; i < OR_LOG_MAX_DEGREE - logn; ++i) phi = or_mulmod(phi, phi, cm);  /* :39-43 */
  uint64_t t = 1;
  for (size_t i = 0; i < degree; ++i) {            /* :47-53 */
    phi_tbl[i] = t;
    phi_shoup[i] = or_shoupify(t, cm);
    t = or_mulmod(t, phi, cm);
  }
  const uint64_t invphi = or_mulmod(t, phi_tbl[degree - 1], cm);  /* :58 */
  t = 1;
  for (size_t i = 0; i < degree; ++i) {            /* :60-65 */
    invphi_tbl[i] = t;
    invphi_shoup[i] = or_shoupify(t, cm);
    t = or_mulmod(t, invphi, cm);
  }
  revbin_permute(phi_tbl, degree);                 /* :68-71 */
  revbin_permute(phi_shoup, degree);
  revbin_permute(invphi_tbl, degree);
  revbin_permute(invphi_shoup, degree);
  scalars[0] = inv_mod_prime(degree, cm);          /* :73-76 */
  scalars[2] = or_mulmod(scalars[0], invphi_tbl[1], cm);
  scalars[1] = or_shoupify(scalars[0], cm);
  scalars[3] = or_shoupify(scalars[2], cm);
  return 0;
}

typedef struct {
  uint64_t *phi, *phi_s, *iphi, *iphi_s;
  uint64_t sc[4];
} or_tab;

static int tab_make(or_tab *t, size_t degree, size_t cm)
{
  uint64_t *buf = (uint64_t *)malloc(4 * degree * sizeof(uint64_t));
  if (!buf) return -1;
  t->phi = buf;
  t->phi_s = buf + degree;
  t->iphi = buf + 2 * degree;
  t->iphi_s = buf + 3 * degree;
  if (or_tables_init(degree, cm, t->phi, t->phi_s, t->iphi, t->iphi_s, t->sc)) {
    free(buf);
    return -1;
  }
  return 0;
}

static void tab_free(or_tab *t) { free(t->phi); }

/* include/yell/details/ntt.hpp:125-134 */
int or_forward_lazy(uint64_t *x, size_t degree, size_t cm)
{
  or_tab t;
  if (tab_make(&t, degree, cm)) return -1;
  int rc = or_negacyclic_forward_lazy(x, degree, t.phi, t.phi_s, kP[cm]);
  tab_free(&t);
  return rc;
}

/* include/yell/details/ntt.hpp:106-123: forward_lazy, then [0,4p) -> [0,p). */
int or_forward(uint64_t *x, size_t degree, size_t cm)
{
  int rc = or_forward_lazy(x, degree, cm);
  if (rc) return rc;
  const uint64_t p = kP[cm];
  for (size_t i = 0; i < degree; ++i) {
    uint64_t v = x[i];
    if (v >= 2 * p) v -= 2 * p;
    if (v >= p) v -= p;
    x[i] = v;
  }
  return 0;
}

/* include/yell/details/ntt.hpp:136-166: backward_lazy, then the merged last
 * layer x0 = (u+v)*N^-1, x1 = (u-v+2p)*N^-1*invphi[1], both by Shoup. */
int or_backward(uint64_t *x, size_t degree, size_t cm)
{
  or_tab t;
  if (degree < 4 || tab_make(&t, degree, cm)) return -1;
  const uint64_t p = kP[cm];
  int rc = or_negacyclic_backward_lazy(x, degree, t.iphi, t.iphi_s, p);
  const size_t half = degree / 2;
  for (size_t i = 0; i < half; ++i) {
    const uint64_t u = x[i], v = x[i + half];
    x[i] = or_mulmod_shoup(u + v, t.sc[0], t.sc[1], cm);
    x[i + half] = or_mulmod_shoup(u - v + 2 * p, t.sc[2], t.sc[3], cm);
  }
  tab_free(&t);
  return rc;
}

int or_poly_forward(uint64_t *x, size_t degree, size_t nmoduli)
{
  for (size_t cm = 0; cm < nmoduli; ++cm)
    if (or_forward(x + cm * degree, degree, cm)) return -1;
  return 0;
}

int or_poly_forward_lazy(uint64_t *x, size_t degree, size_t nmoduli)
{
  for (size_t cm = 0; cm < nmoduli; ++cm)
    if (or_forward_lazy(x + cm * degree, degree, cm)) return -1;
  return 0;
}

int or_poly_backward(uint64_t *x, size_t degree, size_t nmoduli)
{
  for (size_t cm = 0; cm < nmoduli; ++cm)
    if (or_backward(x + cm * degree, degree, cm)) return -1;
  return 0;
}

void or_poly_mul(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli)
{
  for (size_t cm = 0; cm < nmoduli; ++cm)
    for (size_t i = 0; i < degree; ++i)
      out[cm * degree + i] = or_mulmod(a[cm * degree + i], b[cm * degree + i], cm);
}

void or_poly_add(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli)
{
  for (size_t cm = 0; cm < nmoduli; ++cm)
    for (size_t i = 0; i < degree; ++i)
      out[cm * degree + i] = or_addmod(a[cm * degree + i], b[cm * degree + i], cm);
}

void or_poly_sub(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli)
{
  for (size_t cm = 0; cm < nmoduli; ++cm)
    for (size_t i = 0; i < degree; ++i)
      out[cm * degree + i] = or_submod(a[cm * degree + i], b[cm * degree + i], cm);
}

/* include/yell/details/poly.hpp:151-158: v -> p - v (p for v == 0). */
void or_poly_negate(uint64_t *x, size_t degree, size_t nmoduli)
{
  for (size_t cm = 0; cm < nmoduli; ++cm)
    for (size_t i = 0; i < degree; ++i) x[cm * degree + i] = kP[cm] - x[cm * degree + i];
}

/* include/yell/details/poly.hpp:119-132 */
void or_poly_add_product_of(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t degree,
                            size_t nmoduli)
{
  for (size_t cm = 0; cm < nmoduli; ++cm)
    for (size_t i = 0; i < degree; ++i) {
      const size_t k = cm * degree + i;
      acc[k] = or_muladd(acc[k], a[k], b[k], cm);
    }
}

int or_polymul(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli)
{
  const size_t n = degree * nmoduli;
  uint64_t *fa = (uint64_t *)malloc(n * sizeof(uint64_t));
  uint64_t *fb = (uint64_t *)malloc(n * sizeof(uint64_t));
  int rc = -1;
  if (!fa || !fb) goto done;
  memcpy(fa, a, n * sizeof(uint64_t));
  memcpy(fb, b, n * sizeof(uint64_t));
  if (or_poly_forward(fa, degree, nmoduli) || or_poly_forward(fb, degree, nmoduli)) goto done;
  or_poly_mul(c, fa, fb, degree, nmoduli);
  rc = or_poly_backward(c, degree, nmoduli);
done:
  free(fa);
  free(fb);
  return rc;
}

void or_schoolbook(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree, size_t cm)
{
  const uint64_t p = kP[cm];
  for (size_t k = 0; k < degree; ++k) {
    or_u128 pos = 0, neg = 0;
    for (size_t i = 0; i <= k; ++i) pos = (pos + (or_u128)a[i] * b[k - i]) % p;
    for (size_t i = k + 1; i < degree; ++i) neg = (neg + (or_u128)a[i] * b[degree + k - i]) % p;
    c[k] = (uint64_t)((pos + p - neg) % p);
  }
}
