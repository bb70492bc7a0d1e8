/*
 * yell_oracle32.h -- CPU oracle of the reference's 32-bit moduli build
 * (YELL_USE_32BITS_MODULI: include/yell/params_32.hpp, value_type uint32_t,
 * gt_value_type uint64_t, 291 primes of 30 bits, degrees up to 2^15).
 *
 * TEST INFRASTRUCTURE ONLY.  The same restatement as yell_oracle.h (the
 * source yell_oracle.c compiled with -DOR_WORD32); every function cites the
 * reference file:line it follows there.  Pinned (tests/test_oracle32.py) by
 * the reference's own sources compiled in the 32-bit build
 * (oracle/_ref/libyellref32.so), by the closed forms of yell_oracle.h and by
 * a 64-bit-exact schoolbook.
 */
#ifndef YELL_ORACLE32_H
#define YELL_ORACLE32_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

size_t   or32_num_moduli(void);
uint32_t or32_prime(size_t cm);
uint32_t or32_barrett_const(size_t cm);
uint32_t or32_root(size_t cm);

uint32_t or32_barrett(uint64_t x, size_t cm);
uint32_t or32_barrett_hl(uint32_t hi, uint32_t lo, size_t cm);
uint32_t or32_shoupify(uint32_t x, size_t cm);
uint32_t or32_mulmod(uint32_t x, uint32_t y, size_t cm);
uint32_t or32_mulmod_shoup(uint32_t x, uint32_t y, uint32_t yprime, size_t cm);
uint32_t or32_addmod(uint32_t x, uint32_t y, size_t cm);
uint32_t or32_submod(uint32_t x, uint32_t y, size_t cm);
uint32_t or32_muladd(uint32_t rop, uint32_t x, uint32_t y, size_t cm);

void or32_ct_butterfly(uint32_t *x0, uint32_t *x1, uint32_t w, uint32_t wprime, uint32_t p);
void or32_gs_butterfly(uint32_t *x0, uint32_t *x1, uint32_t w, uint32_t wprime, uint32_t p);

int or32_negacyclic_forward_lazy(uint32_t *x, size_t degree, const uint32_t *wtab,
                                 const uint32_t *wtab_shoup, uint32_t p);
int or32_negacyclic_backward_lazy(uint32_t *x, size_t degree, const uint32_t *wtab,
                                  const uint32_t *wtab_shoup, uint32_t p);

int or32_tables_init(size_t degree, size_t cm, uint32_t *phi, uint32_t *phi_shoup,
                     uint32_t *invphi, uint32_t *invphi_shoup, uint32_t *scalars);

int or32_forward(uint32_t *x, size_t degree, size_t cm);
int or32_forward_lazy(uint32_t *x, size_t degree, size_t cm);
int or32_backward(uint32_t *x, size_t degree, size_t cm);

int or32_poly_forward(uint32_t *x, size_t degree, size_t nmoduli);
int or32_poly_forward_lazy(uint32_t *x, size_t degree, size_t nmoduli);
int or32_poly_backward(uint32_t *x, size_t degree, size_t nmoduli);
void or32_poly_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli);
void or32_poly_add(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli);
void or32_poly_sub(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli);
void or32_poly_negate(uint32_t *x, size_t degree, size_t nmoduli);
void or32_poly_add_product_of(uint32_t *acc, const uint32_t *a, const uint32_t *b, size_t degree,
                              size_t nmoduli);
int or32_polymul(uint32_t *c, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli);
void or32_schoolbook(uint32_t *c, const uint32_t *a, const uint32_t *b, size_t degree, size_t cm);

#ifdef __cplusplus
}
#endif
#endif
