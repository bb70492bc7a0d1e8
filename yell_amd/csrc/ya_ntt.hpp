// ya_ntt.hpp -- register/LDS building blocks of the negacyclic NTT on gfx950.
//
// Index conventions.  A residue has N = 2^logn words.  The reference's layer
// with butterfly half-distance h = 2^b (lib/ntt.cpp:102-147 forward, :162-208
// inverse) pairs the words whose indices differ only in bit b; the pair
// (i, i + h) uses twiddle tab[(N >> (b+1)) + (i >> (b+1))] in both
// directions (tab = the bit-reversed phi or inverse-phi table,
// include/yell/details/ntt.hpp:47-71).  The forward transform visits
// b = logn-1 ... 0 with the CT butterfly; the inverse visits b = 0 ... logn-2
// with the GS butterfly and ends with the merged (b = logn-1, x N^-1) layer
// (include/yell/details/ntt.hpp:151-159).  The pairs of one layer are
// disjoint, so any schedule that keeps this per-pair layer order computes
// the same words: that is what lets several layers run in registers.
//
// A "pass" holds E = 2^LE words per thread whose indices differ only in the
// window of bits [WLO, WLO + LE); the remaining index bits come from the
// thread number t: the low WLO bits of i are the low WLO bits of t, the bits
// above the window are t >> WLO.  Between passes the words are exchanged
// through LDS with a bank-spreading layout (lds_slot) so that strided windows
// do not all land in the same banks.
#pragma once

#include "ya_arith.hpp"

namespace ya {

// Per-(degree, prime) device descriptor.  fwd/inv point at N twiddle pairs in
// bit-reversed order (the reference's phiTbl/shoupPhiTbl and
// invphiTbl/shoupInvphiTbl, interleaved).
struct ModDesc {
  const Tw *fwd;
  const Tw *inv;
  uint64_t p;
  uint64_t pn;           // Barrett constant low64(floor(2^128 / p))
  uint64_t inv_n;        // N^-1 mod p and its Shoup companion
  uint64_t inv_n_s;
  uint64_t inv_n_phi;    // N^-1 * invphiTbl[1] and its Shoup companion
  uint64_t inv_n_phi_s;
};

// The same descriptor for the 32-bit moduli build (uint32 words, Tw32 pairs).
struct ModDesc32 {
  const Tw32 *fwd;
  const Tw32 *inv;
  uint32_t p;
  uint32_t pn;           // Barrett constant low32(floor(2^64 / p))
  uint32_t inv_n;
  uint32_t inv_n_s;
  uint32_t inv_n_phi;
  uint32_t inv_n_phi_s;
};

__device__ __forceinline__ uint32_t lds_pad(uint32_t i) { return i + (i >> 4); }

// LDS slot of word i in an exchange buffer.  64-bit words: an XOR swizzle of
// the low five index bits with the five above bit 3 (radix-16 windows) or
// bit 4 (radix-32), which spreads every exchange window of a pass over all
// 64 four-byte banks (the padded layout left 2-way conflicts on the windows
// at bits 6, 8 and 10, a quarter of the LDS cycles).  It only permutes index
// bits 0..4 within a block, so it stays inside [0, 2^LB).  32-bit words keep
// the one-word-in-16 padding.
template <int LE, class W>
__device__ __forceinline__ uint32_t lds_slot(uint32_t i) {
  if constexpr (sizeof(W) == 8) return i ^ ((i >> (LE >= 5 ? 5 : 4)) & 31u);
  return lds_pad(i);
}

// Twiddle load through a global-address-space pointer (the descriptor's
// pointers are generic; without the cast the loads become FLAT and also
// count against the LDS queue).
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ Tw load_tw(const Tw *p) {
  const u64x2 v = *(const __attribute__((address_space(1))) u64x2 *)p;
  return Tw{v.x, v.y};
}
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ Tw32 load_tw(const Tw32 *p) {
  const u32x2 v = *(const __attribute__((address_space(1))) u32x2 *)p;
  return Tw32{v.x, v.y};
}

// A twiddle table read through the constant address space: its wave-uniform
// pairs come through scalar loads.  The persistent batch kernel reads every
// body's twiddles so (there the compiler cannot prove on its own that the
// table is not written before a load -- the previous work item's stores to c
// come first in the loop -- and through the global address space every lane
// would load the same pair on the vector memory path), and so does the fused
// poly-multiplication kernel's dual form (2.2% faster at N = 2^14, L = 8 once
// most of its exchanges waited for their wave only; 2.4% slower while every
// exchange waited for the workgroup, the scalar loads sharing the lgkm counter
// with the LDS exchanges).  The block, split and pair kernels keep the plain
// Tw tables (the same change measured 1-4% slower there).  TwC32 is the same
// for the 32-bit moduli build's fused kernel (0.405 -> 0.395-0.401 ms per
// 256 polys of 2^14 x 16 primes).
struct TwC : Tw {};
template <int WLO>
__device__ __forceinline__ Tw load_tw_w(const Tw *p) {
  return load_tw(p);
}
template <int WLO>
__device__ __forceinline__ Tw32 load_tw_w(const Tw32 *p) {
  return load_tw(p);
}
struct TwC32 : Tw32 {};
template <int WLO>
__device__ __forceinline__ TwC32 load_tw_w(const TwC32 *p) {
  TwC32 r;
  if constexpr (WLO >= 6) {
    const u32x2 v = *(const __attribute__((address_space(4))) u32x2 *)p;
    r.w = v.x;
    r.ws = v.y;
  } else {
    const u32x2 v = *(const __attribute__((address_space(1))) u32x2 *)p;
    r.w = v.x;
    r.ws = v.y;
  }
  return r;
}
template <int WLO>
__device__ __forceinline__ TwC load_tw_w(const TwC *p) {
  TwC r;
  if constexpr (WLO >= 6) {
    const u64x2 v = *(const __attribute__((address_space(4))) u64x2 *)p;
    r.w = v.x;
    r.ws = v.y;
  } else {
    const u64x2 v = *(const __attribute__((address_space(1))) u64x2 *)p;
    r.w = v.x;
    r.ws = v.y;
  }
  return r;
}

// Twiddle base for a pass whose window starts at WLO: the bits of gbase are
// thread bits >= WLO, so for WLO >= 6 they are equal across a 64-lane wave and
// the value is made explicitly wave-uniform (scalar twiddle loads).
template <int WLO>
__device__ __forceinline__ uint32_t uniform_if_wave(uint32_t v) {
  if constexpr (WLO >= 6) return __builtin_amdgcn_readfirstlane(v);
  return v;
}

// Phase isolation.  A fused kernel runs several transforms that load the
// same twiddles and compute the same LDS / global addresses; the optimizer
// would keep those values live from one transform to the next (exceeding the
// 128 registers of a 1024-thread workgroup).  Passing the thread index and a
// zero offset through an empty asm statement makes each phase recompute them.
__device__ __forceinline__ uint32_t phase_v(uint32_t v) {
  __asm__ volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ uint32_t phase_zero_s() {
  uint32_t z = 0;
  __asm__ volatile("" : "+s"(z));
  return z;
}

template <int LB>
struct LdsWords {
  static constexpr uint32_t value = (1u << LB) + ((1u << LB) >> 4);
};

// Global index of word e of thread t for window [WLO, WLO+LE), in a block of
// 2^LB words: i = tlo | e << WLO | thi << (WLO+LE).
template <int WLO, int LE>
__device__ __forceinline__ uint32_t win_base(uint32_t t) {
  const uint32_t tlo = t & ((1u << WLO) - 1u);
  const uint32_t thi = t >> WLO;
  return tlo | (thi << (WLO + LE));
}

// The window helpers below are generic in the word type W (uint64_t, or
// uint32_t in the 32-bit moduli build), deduced from the register array.
template <int WLO, int LE, class W>
__device__ __forceinline__ void lds_store(W *s, const W (&x)[1 << LE], uint32_t t) {
  const uint32_t base = win_base<WLO, LE>(t);
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) s[lds_slot<LE, W>(base | ((uint32_t)e << WLO))] = x[e];
}

template <int WLO, int LE, class W>
__device__ __forceinline__ void lds_load(const W *s, W (&x)[1 << LE], uint32_t t) {
  const uint32_t base = win_base<WLO, LE>(t);
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = s[lds_slot<LE, W>(base | ((uint32_t)e << WLO))];
}

// Global-memory word accesses of a window.  The pointers are cast to the
// global address space (residue pointers read from a descriptor are generic,
// and generic FLAT accesses need a 64-bit address pair per access), and the
// byte offset is formed in 32 bits (a block is at most 2^17 words), so every
// access is a base-register + 32-bit-offset global load/store.
typedef __attribute__((address_space(1))) uint64_t gword;
typedef __attribute__((address_space(1))) unsigned char gbyte;

template <class W>
struct WordShift {
  static constexpr int value = sizeof(W) == 8 ? 3 : 2;
};

template <int WLO, int LE, class W>
__device__ __forceinline__ void gmem_load(const typename Ident<W>::type *__restrict__ g, W (&x)[1 << LE],
                                          uint32_t t) {
  typedef __attribute__((address_space(1))) W gw;
  const gbyte *gb = (const gbyte *)g;
  const uint32_t base = win_base<WLO, LE>(t);
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e)
    x[e] = *(const gw *)(gb + ((base | ((uint32_t)e << WLO)) << WordShift<W>::value));
}

template <int WLO, int LE, class W>
__device__ __forceinline__ void gmem_store(typename Ident<W>::type *__restrict__ g, const W (&x)[1 << LE],
                                           uint32_t t) {
  typedef __attribute__((address_space(1))) W gw;
  gbyte *gb = (gbyte *)g;
  const uint32_t base = win_base<WLO, LE>(t);
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e)
    *(gw *)(gb + ((base | ((uint32_t)e << WLO)) << WordShift<W>::value)) = x[e];
}

// Descriptor read through the constant address space: when the address is
// wave-uniform the fields land in scalar registers (s_load).
__device__ __forceinline__ ModDesc load_desc(const ModDesc *p) {
  const __attribute__((address_space(4))) u64x2 *q = (const __attribute__((address_space(4))) u64x2 *)p;
  const u64x2 v0 = q[0], v1 = q[1], v2 = q[2], v3 = q[3];
  ModDesc d;
  d.fwd = (const Tw *)v0.x;
  d.inv = (const Tw *)v0.y;
  d.p = v1.x;
  d.pn = v1.y;
  d.inv_n = v2.x;
  d.inv_n_s = v2.y;
  d.inv_n_phi = v3.x;
  d.inv_n_phi_s = v3.y;
  return d;
}

// Forward (CT) layers b = HI-1 ... LO of one pass.  `gbase` is the residue
// index of word 0 of this thread with the window bits and the bits below the
// window cleared (i.e. block offset | thi << (WLO+LE)); nlog = log2 N of the
// whole residue.
template <int WLO, int LE, int HI, int LO, class W, class TW>
__device__ __forceinline__ void fwd_layers(W (&x)[1 << LE], const TW *__restrict__ tw, uint32_t nlog,
                                           uint32_t gbase, typename Ident<W>::type p) {
  const typename WordK<W>::K k(p);
#pragma unroll
  for (int b = HI - 1; b >= LO; --b) {
    const int j = b - WLO;
    const uint32_t toff = ((1u << nlog) >> (b + 1)) + (uniform_if_wave<WLO>(gbase) >> (b + 1));
#pragma unroll
    for (int g = 0; g < (1 << (LE - 1 - j)); ++g) {
      const TW t = load_tw_w<WLO>(tw + toff + g);
#pragma unroll
      for (int q = 0; q < (1 << j); ++q) {
        const int e0 = (g << (j + 1)) | q;
        ct_butterfly(x[e0], x[e0 | (1 << j)], t, k);
      }
    }
  }
}

// The same forward layers applied to two residues of one prime at once (the
// two operands of a poly-multiplication): each twiddle is loaded once and
// used for both, and the two butterfly chains are independent.
template <int WLO, int LE, int HI, int LO, class W, class TW>
__device__ __forceinline__ void fwd_layers2(W (&x)[1 << LE], W (&y)[1 << LE], const TW *__restrict__ tw,
                                            uint32_t nlog, uint32_t gbase, typename Ident<W>::type p) {
  const typename WordK<W>::K k(p);
#pragma unroll
  for (int b = HI - 1; b >= LO; --b) {
    const int j = b - WLO;
    const uint32_t toff = ((1u << nlog) >> (b + 1)) + (uniform_if_wave<WLO>(gbase) >> (b + 1));
#pragma unroll
    for (int g = 0; g < (1 << (LE - 1 - j)); ++g) {
      const TW t = load_tw_w<WLO>(tw + toff + g);
#pragma unroll
      for (int q = 0; q < (1 << j); ++q) {
        const int e0 = (g << (j + 1)) | q;
        ct_butterfly(x[e0], x[e0 | (1 << j)], t, k);
        ct_butterfly(y[e0], y[e0 | (1 << j)], t, k);
      }
    }
  }
}

// Inverse (GS) layers b = LO ... HI-1 of one pass (same indexing).
template <int WLO, int LE, int HI, int LO, class W, class TW>
__device__ __forceinline__ void inv_layers(W (&x)[1 << LE], const TW *__restrict__ tw, uint32_t nlog,
                                           uint32_t gbase, typename Ident<W>::type p) {
  const typename WordK<W>::K k(p);
#pragma unroll
  for (int b = LO; b < HI; ++b) {
    const int j = b - WLO;
    const uint32_t toff = ((1u << nlog) >> (b + 1)) + (uniform_if_wave<WLO>(gbase) >> (b + 1));
#pragma unroll
    for (int g = 0; g < (1 << (LE - 1 - j)); ++g) {
      const TW t = load_tw_w<WLO>(tw + toff + g);
#pragma unroll
      for (int q = 0; q < (1 << j); ++q) {
        const int e0 = (g << (j + 1)) | q;
        gs_butterfly(x[e0], x[e0 | (1 << j)], t, k);
      }
    }
  }
}

// The merged last inverse layer on window bit J (include/yell/details/ntt.hpp:151-159):
// x0 = (u + v) * N^-1, x1 = (u - v + 2p) * N^-1 * invphi[1], each by Shoup, canonical.
template <int LE, int J, class W, class D>
__device__ __forceinline__ void inv_last_layer(W (&x)[1 << LE], const D &d) {
  const W p = d.p, p2 = p << 1;
#pragma unroll
  for (int e0 = 0; e0 < (1 << LE); ++e0) {
    if (e0 & (1 << J)) continue;
    const W u = x[e0], v = x[e0 | (1 << J)];
    x[e0] = mulmod_shoup(u + v, d.inv_n, d.inv_n_s, p);
    x[e0 | (1 << J)] = mulmod_shoup(u - v + p2, d.inv_n_phi, d.inv_n_phi_s, p);
  }
}

template <int LE, class W>
__device__ __forceinline__ void correct_all(W (&x)[1 << LE], typename Ident<W>::type p) {
  const W p2 = p << 1;
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = correct4p(x[e], p, p2);
}

}  // namespace ya
