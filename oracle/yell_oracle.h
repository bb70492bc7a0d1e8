/*
 * yell_oracle.h -- CPU oracle for the negacyclic NTT / RNS arithmetic path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * arithmetic (fionser/YELL), used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker.  The product path (yell_amd/)
 * never links, loads or calls anything in this directory.
 *
 * Each function cites the reference file:line it restates.  Values are
 * uint64_t residues modulo 62-bit primes P[cm]; u128 is the compiler's
 * 128-bit unsigned integer, as in params_64.hpp:8-11 of the reference.
 *
 * Parity pinning (see tests/test_oracle.py):
 *   - or_barrett / or_mulmod / or_muladd  : against Python big-int x mod p;
 *   - or_mulmod_shoup / or_shoupify       : against floor(w*2^64/p) and x*w mod p;
 *   - or_tables_init                      : phi^N == -1, phi^(2N) == 1, bit-reversal,
 *                                           N * N^-1 == 1, Shoup companions;
 *   - or_forward                          : closed form A[j] = sum a_i phi^((2brv(j)+1) i);
 *   - or_forward_lazy                     : the compiled reference (oracle/_ref) on
 *                                           the same inputs, where it builds;
 *   - or_backward                         : backward(forward(a)) == a and the compiled
 *                                           reference;
 *   - or_polymul                          : negacyclic schoolbook convolution
 *                                           (or_schoolbook, Python ints, GMP).
 */
#ifndef YELL_ORACLE_H
#define YELL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned __int128 or_u128;

/* Constant tables (generated, see yell_amd/csrc/gen_params.py). */
size_t   or_num_moduli(void);
uint64_t or_prime(size_t cm);
uint64_t or_barrett_const(size_t cm);
uint64_t or_root(size_t cm);

/* Scalar modular operations (include/yell/ops.hpp:46-158, lib/ops.cpp:5-27). */
uint64_t or_barrett(or_u128 x, size_t cm);
uint64_t or_barrett_hl(uint64_t hi, uint64_t lo, size_t cm); /* ctypes-friendly */
uint64_t or_shoupify(uint64_t x, size_t cm);
uint64_t or_mulmod(uint64_t x, uint64_t y, size_t cm);
uint64_t or_mulmod_shoup(uint64_t x, uint64_t y, uint64_t yprime, size_t cm);
uint64_t or_addmod(uint64_t x, uint64_t y, size_t cm);
uint64_t or_submod(uint64_t x, uint64_t y, size_t cm);
uint64_t or_muladd(uint64_t rop, uint64_t x, uint64_t y, size_t cm);

/* Butterflies (lib/ntt.cpp:47-83).  In/out through pointers, as the reference. */
void or_ct_butterfly(uint64_t *x0, uint64_t *x1, uint64_t w, uint64_t wprime, uint64_t p);
void or_gs_butterfly(uint64_t *x0, uint64_t *x1, uint64_t w, uint64_t wprime, uint64_t p);

/* The out-of-line NTT loops (lib/ntt.cpp:89-210).  Return 0, or -1 for an
 * unsupported degree (forward needs degree >= 4, backward degree >= 4). */
int or_negacyclic_forward_lazy(uint64_t *x, size_t degree, const uint64_t *wtab,
                               const uint64_t *wtab_shoup, uint64_t p);
int or_negacyclic_backward_lazy(uint64_t *x, size_t degree, const uint64_t *wtab,
                                const uint64_t *wtab_shoup, uint64_t p);

/* Per-(degree, cm) precomputed tables (include/yell/details/ntt.hpp:27-77).
 * phi, phi_shoup, invphi, invphi_shoup: arrays of `degree` words (bit-reversed);
 * scalars[4] = {invDegree, shoupInvDegree, invphiInvDegree, shoupInvphiInvDegree}. */
int or_tables_init(size_t degree, size_t cm, uint64_t *phi, uint64_t *phi_shoup,
                   uint64_t *invphi, uint64_t *invphi_shoup, uint64_t *scalars);

/* ntt<N>::forward / forward_lazy / backward on one residue
 * (include/yell/details/ntt.hpp:106-166).  Tables are built on each call. */
int or_forward(uint64_t *x, size_t degree, size_t cm);
int or_forward_lazy(uint64_t *x, size_t degree, size_t cm);
int or_backward(uint64_t *x, size_t degree, size_t cm);

/* Residue-major [nmoduli][degree] poly operations; residue cm uses P[cm]
 * (include/yell/ops.hpp:7-41, include/yell/details/poly.hpp:119-186). */
int or_poly_forward(uint64_t *x, size_t degree, size_t nmoduli);
int or_poly_forward_lazy(uint64_t *x, size_t degree, size_t nmoduli);
int or_poly_backward(uint64_t *x, size_t degree, size_t nmoduli);
void or_poly_mul(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli);
void or_poly_add(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli);
void or_poly_sub(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli);
void or_poly_negate(uint64_t *x, size_t degree, size_t nmoduli);
void or_poly_add_product_of(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t degree,
                            size_t nmoduli);

/* One RNS poly-multiplication: c = backward(forward(a) * forward(b)), the
 * reference call sequence of details/poly.hpp:171-186 + ops.hpp:7-24. */
int or_polymul(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli);

/* Textbook negacyclic convolution mod P[cm] with 128-bit accumulation:
 * c_k = sum_{i+j=k} a_i b_j - sum_{i+j=k+N} a_i b_j  (mod p).  O(N^2). */
void or_schoolbook(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree, size_t cm);

#ifdef __cplusplus
}
#endif
#endif
