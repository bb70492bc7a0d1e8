// ya_arith.hpp -- 64-bit modular arithmetic for gfx950 built from 32-bit
// VALU multiplies (CDNA4 has no 64x64-bit multiply instruction).
//
// Every routine returns exactly the word the reference's scalar code returns:
//   barrett_reduce   include/yell/ops.hpp:85-99 + lib/ops.cpp:5-21
//   mulmod_shoup     include/yell/ops.hpp:122-142
//   ct_butterfly     lib/ntt.cpp:69-83   (lazy, [0,4p) -> [0,4p))
//   gs_butterfly     lib/ntt.cpp:47-63   (lazy, [0,2p) -> [0,2p))
//   addmod / submod  include/yell/ops.hpp:54-83
// The primes are < 2^62, so 2p and 4p - 1 fit in a 64-bit word.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ya {

// A twiddle factor with its Shoup companion w' = floor(w * 2^64 / p).
struct __attribute__((aligned(16))) Tw {
  uint64_t w, ws;
};

__device__ __forceinline__ uint32_t lo32(uint64_t v) { return (uint32_t)v; }
__device__ __forceinline__ uint32_t hi32(uint64_t v) { return (uint32_t)(v >> 32); }

// Full 32x32 -> 64 product (one v_mad_u64_u32 with a zero addend).
__device__ __forceinline__ uint64_t mul32x32(uint32_t a, uint32_t b) {
  return (uint64_t)a * (uint64_t)b;
}

// ---- Short assembly sequences.  The compiler neither exposes the
// carry-out of v_mad_u64_u32 nor chains a product into a full 64-bit addend
// (it zero-extends 32-bit halves into fresh register pairs, one v_mov each),
// and it rewrites x >= m ? x - m : x into a compare, two masked selects of m
// and a borrow chain.  Each instruction reads at most one scalar operand (the
// VOP3 constant-bus limit).  hipcc places an operand constrained "s" in a
// scalar register where the value is wave-uniform and, where it is not (the
// persistent batch kernel's sub-wave rings, whose lanes may hold different
// primes), in a vector register -- both legal VOP3 sources; the batch tests
// over mixed rings check those paths word for word.  The sequences write no
// memory and read nothing but their operands.

// a * b + c (mod 2^64) and the carry out of bit 63 (0 or 1).  The select
// that reads the carry mask waits two cycles after the multiply wrote it
// (VALU-written SGPR read as a lane mask).
__device__ __forceinline__ uint64_t mad64_carry(uint32_t a, uint32_t b, uint64_t c, uint32_t &carry) {
  uint64_t d, mask;
  asm("v_mad_u64_u32 %0, %2, %3, %4, %5\n\ts_nop 1\n\tv_cndmask_b32_e64 %1, 0, 1, %2"
      : "=v"(d), "=v"(carry), "=s"(mask)
      : "v"(a), "s"(b), "v"(c));
  return d;
}

// a * b + c (mod 2^64) with c read as one 64-bit register pair.
__device__ __forceinline__ uint64_t mad64_wrap(uint32_t a, uint32_t b, uint64_t c) {
  uint64_t mask;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(c), "=s"(mask) : "v"(a), "s"(b));
  return c;
}

// x >= m ? x - m : x, with nm = 2^64 - m: the difference as one opaque 64-bit
// add, then one compare and two selects.
__device__ __forceinline__ uint64_t sub_if_ge(uint64_t x, uint64_t m, uint64_t nm) {
  uint64_t d;
  asm("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(d) : "v"(x), "s"(nm));
  return x >= m ? d : x;
}

// High 64 bits of the 128-bit product a * b: the middle column
// a0*b1 + a1*b0 + hi32(a0*b0) is up to 2^65, so its carry out of bit 63 is
// kept and added at bit 32 of the top product.
__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
  const uint32_t a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
  const uint64_t m1 = mul32x32(a1, b0) + (uint64_t)__umulhi(a0, b0);  // < 2^64
  uint32_t c;
  const uint64_t m2 = mad64_carry(a0, b1, m1, c);
  return mul32x32(a1, b1) + (((uint64_t)c << 32) | hi32(m2));
}

// Low 64 bits of a * b.
__device__ __forceinline__ uint64_t mullo64(uint64_t a, uint64_t b) {
  const uint32_t a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
  const uint64_t p00 = mul32x32(a0, b0);
  const uint32_t cross = a0 * b1 + a1 * b0;
  return p00 + ((uint64_t)cross << 32);
}

// r with its high word replaced by h, written on the two 32-bit halves: the
// equivalent (h << 32) | lo32(r), with h = hi32(r) + s, is rewritten by the
// optimizer into the 64-bit add r + (s << 32), a zero-extension move and a
// 64-bit add where one 32-bit add suffices.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint64_t with_hi(uint64_t r, uint32_t h) {
  u32x2 v = __builtin_bit_cast(u32x2, r);
  v.y = h;
  return __builtin_bit_cast(uint64_t, v);
}

// 128-bit product as (hi, lo), the middle column as in mulhi64.
__device__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
  const uint32_t a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
  const uint64_t p00 = mul32x32(a0, b0);
  const uint64_t m1 = mul32x32(a1, b0) + hi32(p00);
  uint32_t c;
  const uint64_t m2 = mad64_carry(a0, b1, m1, c);
  hi = mul32x32(a1, b1) + (((uint64_t)c << 32) | hi32(m2));
  lo = with_hi(p00, lo32(m2));
}

// a * b + c with only the low 32 bits of the result wanted: one
// v_mad_u64_u32 whose addend is the previous step's register pair (its high
// half is don't-care), so a sum of 32-bit products is a chain of multiply-adds
// with no separate add.
__device__ __forceinline__ uint64_t mad_lo_acc(uint32_t a, uint32_t b, uint64_t c) {
  uint64_t d, mask;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(d), "=s"(mask) : "v"(a), "s"(b), "v"(c));
  return d;
}

// a * b + c (mod 2^64) into a register pair of its own (c stays live).
__device__ __forceinline__ uint64_t mad64_to(uint32_t a, uint32_t b, uint64_t c) {
  uint64_t d, mask;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(d), "=s"(mask) : "v"(a), "s"(b), "v"(c));
  return d;
}

// 2*x + m (mod 2^64) as one v_lshl_add_u64 (the optimizer otherwise forms
// (x + m/2) << 1 in two instructions).
__device__ __forceinline__ uint64_t dbl_add(uint64_t x, uint64_t m) {
  uint64_t d;
  asm("v_lshl_add_u64 %0, %1, 1, %2" : "=v"(d) : "v"(x), "s"(m));
  return d;
}

// acc + (x*w - q*p) (mod 2^64) with np = 2^64 - p: the cross products
// x0*w1 + x1*w0 + q0*n1 + q1*n0 (they only reach the high word) summed by a
// multiply-add chain, the low products acc + x0*w0 + q0*n0 by another, the
// cross sum added to the high word in between.  Seven VALU instructions where
// the low product, four 32-bit multiplies, two three-way adds and the add of
// acc took nine.
__device__ __forceinline__ uint64_t add_mul_sub_lo64(uint64_t acc, uint64_t x, uint64_t w, uint64_t q,
                                                     uint64_t np) {
  const uint32_t x0 = lo32(x), x1 = hi32(x), w0 = lo32(w), w1 = hi32(w);
  const uint32_t q0 = lo32(q), q1 = hi32(q), n0 = lo32(np), n1 = hi32(np);
  uint64_t c = mul32x32(x0, w1);
  c = mad_lo_acc(x1, w0, c);
  c = mad_lo_acc(q0, n1, c);
  c = mad_lo_acc(q1, n0, c);
  const uint64_t a = mad64_to(x0, w0, acc);
  return mad64_wrap(q0, n0, with_hi(a, hi32(a) + lo32(c)));
}

// x*w - q*p (mod 2^64), the same chains without an addend (the low product
// starts the second).
__device__ __forceinline__ uint64_t mul_sub_lo64(uint64_t x, uint64_t w, uint64_t q, uint64_t np) {
  const uint32_t x0 = lo32(x), x1 = hi32(x), w0 = lo32(w), w1 = hi32(w);
  const uint32_t q0 = lo32(q), q1 = hi32(q), n0 = lo32(np), n1 = hi32(np);
  uint64_t c = mul32x32(x0, w1);
  c = mad_lo_acc(x1, w0, c);
  c = mad_lo_acc(q0, n1, c);
  c = mad_lo_acc(q1, n0, c);
  const uint64_t a = mul32x32(x0, w0);
  return mad64_wrap(q0, n0, with_hi(a, hi32(a) + lo32(c)));
}

// acc + q*np (mod 2^64), acc the multiply-add's 64-bit addend.
__device__ __forceinline__ uint64_t add_mul_lo64(uint64_t acc, uint64_t q, uint64_t np) {
  const uint64_t r = mad64_wrap(lo32(q), lo32(np), acc);
  return with_hi(r, hi32(r) + lo32(q) * hi32(np) + hi32(q) * lo32(np));
}

// Reference Barrett reduction of the 128-bit value (hi, lo) modulo p:
//   q = pn * hi + (x << 2)  (mod 2^128);  r = lo - hi64(q) * p  (mod 2^64);
//   r -= p if r >= p.
__device__ __forceinline__ uint64_t barrett_reduce(uint64_t hi, uint64_t lo, uint64_t p,
                                                   uint64_t pn) {
  uint64_t qh, ql;
  mul128(pn, hi, qh, ql);
  const uint64_t sl = lo << 2;
  const uint64_t sh = (hi << 2) | (lo >> 62);
  const uint64_t s_lo = ql + sl;
  const uint64_t q_hi = qh + sh + (s_lo < ql ? 1u : 0u);
  const uint64_t np = 0 - p;
  const uint64_t r = add_mul_lo64(lo, q_hi, np);  // lo - q_hi * p
  return sub_if_ge(r, p, np);
}

// ops::mulmod (Barrett on the full product).
__device__ __forceinline__ uint64_t mulmod_barrett(uint64_t x, uint64_t y, uint64_t p,
                                                   uint64_t pn) {
  uint64_t hi, lo;
  mul128(x, y, hi, lo);
  return barrett_reduce(hi, lo, p, pn);
}

// ops::muladd: (x * y + acc) over 128 bits, then Barrett.
__device__ __forceinline__ uint64_t muladd_barrett(uint64_t acc, uint64_t x, uint64_t y,
                                                   uint64_t p, uint64_t pn) {
  uint64_t hi, lo;
  mul128(x, y, hi, lo);
  const uint64_t l2 = lo + acc;
  hi += (l2 < lo) ? 1u : 0u;
  return barrett_reduce(hi, l2, p, pn);
}

// ops::mulmod_shoup: x < 4p, result in [0, p).
__device__ __forceinline__ uint64_t mulmod_shoup(uint64_t x, uint64_t y, uint64_t ys,
                                                 uint64_t p) {
  const uint64_t np = 0 - p;
  const uint64_t r = mul_sub_lo64(x, y, mulhi64(x, ys), np);
  return sub_if_ge(r, p, np);
}

// Per-prime constants of the butterflies (wave-uniform where the prime is).
struct PrimeK {
  uint64_t p, p2;    // p, 2p
  uint64_t np, np2;  // 2^64 - p, 2^64 - 2p
  __device__ __forceinline__ explicit PrimeK(uint64_t q)
      : p(q), p2(q << 1), np(0 - q), np2(0 - (q << 1)) {}
};

// Shoup product without the final correction: x * w mod p in [0, 2p), the
// word x*w - hi64(x*w')*p (mod 2^64) of the reference.
__device__ __forceinline__ uint64_t mulmod_shoup_lazy(uint64_t x, const Tw &t, const PrimeK &k) {
  return mul_sub_lo64(x, t.w, mulhi64(x, t.ws), k.np);
}

// lib/ntt.cpp:69-83: x0 = u0 + v, x1 = u0 + 2p - v with v the lazy Shoup
// product.  u0 + v comes out of the product's multiply-add chain (u0 as its
// addend), and x1 = 2*u0 + 2p - x0 (mod 2^64), the same word.
__device__ __forceinline__ void ct_butterfly(uint64_t &x0, uint64_t &x1, const Tw &t, const PrimeK &k) {
  const uint64_t u0 = sub_if_ge(x0, k.p2, k.np2);
  const uint64_t s = add_mul_sub_lo64(u0, x1, t.w, mulhi64(x1, t.ws), k.np);
  x0 = s;
  x1 = dbl_add(u0, k.p2) - s;
}

// lib/ntt.cpp:47-63
__device__ __forceinline__ void gs_butterfly(uint64_t &x0, uint64_t &x1, const Tw &t, const PrimeK &k) {
  const uint64_t u0 = x0, u1 = x1;
  const uint64_t t0 = sub_if_ge(u0 + u1, k.p2, k.np2);
  const uint64_t t1 = u0 + k.p2 - u1;
  x0 = t0;
  x1 = mulmod_shoup_lazy(t1, t, k);
}

// include/yell/details/ntt.hpp:111-118: [0,4p) -> [0,p).
__device__ __forceinline__ uint64_t correct4p(uint64_t v, uint64_t p, uint64_t p2) {
  v = v >= p2 ? v - p2 : v;
  return v >= p ? v - p : v;
}

__device__ __forceinline__ uint64_t addmod(uint64_t x, uint64_t y, uint64_t p) {
  const uint64_t r = x + y;
  return r >= p ? r - p : r;
}

__device__ __forceinline__ uint64_t submod(uint64_t x, uint64_t y, uint64_t p) {
  const uint64_t r = x + (p - y);
  return r >= p ? r - p : r;
}

// ---- 32-bit moduli build (YELL_USE_32BITS_MODULI, include/yell/params_32.hpp):
// value_type uint32_t, gt_value_type uint64_t, primes < 2^30, so 4p < 2^32.
// Every product is one native 32-bit multiply (v_mul_lo_u32 / v_mul_hi_u32)
// or one v_mad_u64_u32; in the butterflies the remainder x*w - q*p (mod 2^32)
// is the low word of the chain mad(q, 2^32 - p, mad(x, w, 0)), two
// instructions where two low multiplies and a subtract would be three.

struct __attribute__((aligned(8))) Tw32 {
  uint32_t w, ws;  // w' = floor(w * 2^32 / p)
};

struct PrimeK32 {
  uint32_t p, p2, np;  // p, 2p, 2^32 - p
  __device__ __forceinline__ explicit PrimeK32(uint32_t q) : p(q), p2(q << 1), np(0 - q) {}
};

// x*w - q*p (mod 2^32).
__device__ __forceinline__ uint32_t mul_sub_lo32(uint32_t x, uint32_t w, uint32_t q, uint32_t np) {
  return lo32(mad64_wrap(q, np, mul32x32(x, w)));
}

// lib/ops.cpp:5-21 with shift = 32, delta = 32 - 30 = 2:
// q = Pn * (x >> 32) + (x << 2) (mod 2^64); r = lo32(x) - hi32(q) * p; one -p.
__device__ __forceinline__ uint32_t barrett_reduce32(uint64_t x, uint32_t p, uint32_t pn) {
  const uint64_t q = (uint64_t)pn * (uint32_t)(x >> 32) + (x << 2);
  const uint32_t r = (uint32_t)x - (uint32_t)(q >> 32) * p;
  return r >= p ? r - p : r;
}

__device__ __forceinline__ uint32_t mulmod_barrett(uint32_t x, uint32_t y, uint32_t p, uint32_t pn) {
  return barrett_reduce32((uint64_t)x * y, p, pn);
}

__device__ __forceinline__ uint32_t muladd_barrett(uint32_t acc, uint32_t x, uint32_t y, uint32_t p,
                                                   uint32_t pn) {
  return barrett_reduce32((uint64_t)x * y + acc, p, pn);
}

// include/yell/ops.hpp:122-142 with 32-bit words, x < 4p.
__device__ __forceinline__ uint32_t mulmod_shoup(uint32_t x, uint32_t y, uint32_t ys, uint32_t p) {
  const uint32_t r = x * y - __umulhi(x, ys) * p;
  return r >= p ? r - p : r;
}

// lib/ntt.cpp:69-83 (bit_width = 32)
__device__ __forceinline__ void ct_butterfly(uint32_t &x0, uint32_t &x1, const Tw32 &t, const PrimeK32 &k) {
  const uint32_t u0 = x0 >= k.p2 ? x0 - k.p2 : x0;
  const uint32_t v = mul_sub_lo32(x1, t.w, __umulhi(x1, t.ws), k.np);
  x0 = u0 + v;
  x1 = u0 + k.p2 - v;
}

// lib/ntt.cpp:47-63 (bit_width = 32)
__device__ __forceinline__ void gs_butterfly(uint32_t &x0, uint32_t &x1, const Tw32 &t, const PrimeK32 &k) {
  const uint32_t u0 = x0, u1 = x1;
  uint32_t t0 = u0 + u1;
  t0 = t0 >= k.p2 ? t0 - k.p2 : t0;
  const uint32_t t1 = u0 + k.p2 - u1;
  x0 = t0;
  x1 = mul_sub_lo32(t1, t.w, __umulhi(t1, t.ws), k.np);
}

__device__ __forceinline__ uint32_t correct4p(uint32_t v, uint32_t p, uint32_t p2) {
  v = v >= p2 ? v - p2 : v;
  return v >= p ? v - p : v;
}

__device__ __forceinline__ uint32_t addmod(uint32_t x, uint32_t y, uint32_t p) {
  const uint32_t r = x + y;
  return r >= p ? r - p : r;
}

__device__ __forceinline__ uint32_t submod(uint32_t x, uint32_t y, uint32_t p) {
  const uint32_t r = x + (p - y);
  return r >= p ? r - p : r;
}

// Word-type traits: the per-prime constants of each build.
template <class W>
struct WordK;
template <>
struct WordK<uint64_t> {
  using K = PrimeK;
};
template <>
struct WordK<uint32_t> {
  using K = PrimeK32;
};

template <class T>
struct Ident {
  using type = T;
};

}  // namespace ya
