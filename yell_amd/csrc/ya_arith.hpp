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

// High 64 bits of the 128-bit product a * b.
__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
  const uint32_t a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
  const uint64_t p00h = (uint64_t)__umulhi(a0, b0);
  const uint64_t m1 = mul32x32(a1, b0) + p00h;        // < 2^64
  const uint64_t m2 = mul32x32(a0, b1) + lo32(m1);    // < 2^64
  return mul32x32(a1, b1) + hi32(m1) + hi32(m2);
}

// Low 64 bits of a * b.
__device__ __forceinline__ uint64_t mullo64(uint64_t a, uint64_t b) {
  const uint32_t a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
  const uint64_t p00 = mul32x32(a0, b0);
  const uint32_t cross = a0 * b1 + a1 * b0;
  return p00 + ((uint64_t)cross << 32);
}

// 128-bit product as (hi, lo).
__device__ __forceinline__ void mul128(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
  const uint32_t a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b);
  const uint64_t p00 = mul32x32(a0, b0);
  const uint64_t m1 = mul32x32(a1, b0) + hi32(p00);
  const uint64_t m2 = mul32x32(a0, b1) + lo32(m1);
  hi = mul32x32(a1, b1) + hi32(m1) + hi32(m2);
  lo = ((uint64_t)lo32(m2) << 32) | lo32(p00);
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
  uint64_t r = lo - mullo64(q_hi, p);
  return r >= p ? r - p : r;
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
  const uint64_t q = mulhi64(x, ys);
  const uint64_t r = mullo64(x, y) - mullo64(q, p);
  return r >= p ? r - p : r;
}

// Per-prime constants of the butterflies (wave-uniform, scalar registers).
struct PrimeK {
  uint64_t p, p2;    // p, 2p
  __device__ __forceinline__ explicit PrimeK(uint64_t q) : p(q), p2(q << 1) {}
};

// Shoup product without the final correction: x * w mod p in [0, 2p), the
// word x*w - hi64(x*w')*p (mod 2^64) of the reference.
__device__ __forceinline__ uint64_t mulmod_shoup_lazy(uint64_t x, const Tw &t, const PrimeK &k) {
  const uint64_t q = mulhi64(x, t.ws);
  return mullo64(x, t.w) - mullo64(q, k.p);
}

// lib/ntt.cpp:69-83
__device__ __forceinline__ void ct_butterfly(uint64_t &x0, uint64_t &x1, const Tw &t, const PrimeK &k) {
  const uint64_t u0 = x0 >= k.p2 ? x0 - k.p2 : x0;
  const uint64_t v = mulmod_shoup_lazy(x1, t, k);
  x0 = u0 + v;
  x1 = u0 + k.p2 - v;
}

// lib/ntt.cpp:47-63
__device__ __forceinline__ void gs_butterfly(uint64_t &x0, uint64_t &x1, const Tw &t, const PrimeK &k) {
  const uint64_t u0 = x0, u1 = x1;
  uint64_t t0 = u0 + u1;
  t0 = t0 >= k.p2 ? t0 - k.p2 : t0;
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
// or one v_mad_u64_u32.

struct __attribute__((aligned(8))) Tw32 {
  uint32_t w, ws;  // w' = floor(w * 2^32 / p)
};

struct PrimeK32 {
  uint32_t p, p2;
  __device__ __forceinline__ explicit PrimeK32(uint32_t q) : p(q), p2(q << 1) {}
};

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
  const uint32_t v = x1 * t.w - __umulhi(x1, t.ws) * k.p;
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
  x1 = t1 * t.w - __umulhi(t1, t.ws) * k.p;
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
