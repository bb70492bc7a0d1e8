// ya_blocks.hpp -- workgroup-level NTT building blocks shared by the kernel
// translation units: a forward transform of one block of 2^LB words ending
// in registers, and an inverse transform starting from registers.
//
// Block kernels transform 2^LB <= 2^14 consecutive words of a residue in one
// workgroup: T = 2^(LB-LE) threads hold E = 2^LE words each, run LE layers in
// registers per pass and exchange words through LDS between passes.  For
// N > 2^14 a column ("top") kernel runs the layers on index bits [14, logn)
// first (forward) or last (inverse), so every block still fits one
// workgroup's LDS.
#pragma once

#include <type_traits>

#include "ya_kernels.hpp"

namespace ya {

// Staggered start.  When every workgroup of a launch fills a CU and runs the
// same fused body, the first wave starts together and stays in step: all CUs
// load their operands at once (HBM at its peak rate, no VALU work), then all
// compute while the HBM idles.  The first `first` workgroups (one per CU)
// wait (b / 8 mod 16) * steps * s_sleep(32) (~0.85 us per step at 2.4 GHz)
// before they start -- 16 groups per XCD, since consecutive workgroups go
// to consecutive XCDs -- and each CU's later workgroups keep its offset, so
// one CU's loads overlap other CUs' butterflies.  `word` = first << 8 |
// steps; 0 = no wait.  The wait is bounded and wave-uniform.
__device__ __forceinline__ void stagger_start(uint32_t word) {
  const uint32_t first = word >> 8, steps = word & 255u;
  if (blockIdx.x < first) {
    for (uint32_t k = ((blockIdx.x >> 3) & 15u) * steps; k != 0; --k) __builtin_amdgcn_s_sleep(32);
  }
}

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// A block's base address passed through the scalar unit as an opaque
// wave-uniform value.  When the compiler sees the address as kernel argument
// plus an offset from blockIdx it rebuilds the per-word addresses in vector
// registers, and the 64-bit multiply chains then lose their zero-extended
// register pairs to extra v_mov copies (the fused 2^14 kernel: 11,845 VALU
// instructions per wave instead of 10,534).
template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (T *)(((uint64_t)hi << 32) | lo);
}

// XCD-aware order: workgroups b, b+8, b+16, ... run on one XCD (the hardware
// deals consecutive workgroups to consecutive XCDs); give each XCD a
// contiguous range [xcd*g/8, ...) of the g work units.
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t g) {
  const uint32_t xcd = b % 8u, pos = b / 8u;
  return xcd * (g / 8u) + min(xcd, g % 8u) + pos;
}

// ---------------------------------------------------------------------------
// Barriers between the passes of a block transform.  WgSync is the
// workgroup's barrier; WaveSync orders one wave's LDS accesses, all a
// transform of one wave or less needs (the persistent batch kernel's
// sub-groups of that size, ya_batch.hip).  Each also has local(), enough for
// an exchange in which every wave reads back only words it wrote itself
// (kWaveLocal below): a wave's LDS accesses are performed in issue order, so
// the wave waits for its own LDS operations and the compiler is kept from
// moving LDS accesses across the point.
struct WaveSync {
  __device__ __forceinline__ void operator()() const {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __device__ __forceinline__ void local() const { (*this)(); }
};

struct WgSync {
  __device__ __forceinline__ void operator()() const { __syncthreads(); }
  __device__ __forceinline__ void local() const { WaveSync()(); }
};

// An exchange from a pass over window [WA, WA+LE) to one over [WB, WB+LE)
// of a block of 2^LB words held by 2^(LB-LE) threads, 64 to a wave: with
// the window at bit 6 or below, wave w holds exactly the 2^(6+LE) words
// whose bits [6+LE, LB) are w (its lanes and registers cover bits
// [0, 6+LE)), so if both windows start at or below bit 6 every word a wave
// reads back in the exchange is one it stored.
template <int LB, int LE, int WA, int WB>
constexpr bool kWaveLocal = (LB - LE >= 6) && WA <= 6 && WB <= 6;

template <bool LOCAL, class S>
__device__ __forceinline__ void xsync(S &sync) {
  if constexpr (LOCAL) {
    sync.local();
  } else {
    sync();
  }
}

// Window start of the first forward pass of a 2^LB block.
template <int LB, int LE>
struct FirstWindow {
  static constexpr int value = LB - LE > 0 ? LB - LE : 0;
};

// Forward transform of index bits [0, LB) of one block, words loaded from
// global `src` (block base), or already in x (PRELOADED: loaded with
// gmem_load<FirstWindow<LB,LE>::value, LE>).  On return x holds the words of
// window [0, LE) for thread t (i = e | t << LE), in [0, 4p).
template <int LB, int LE, bool PRELOADED = false, class W, class TW, class S = WgSync>
__device__ __forceinline__ void fwd_block(W (&x)[1 << LE], const typename Ident<W>::type *__restrict__ src,
                                          W *lds, const TW *__restrict__ tw, uint32_t nlog, uint32_t boff,
                                          typename Ident<W>::type p, uint32_t t, S &&sync = S()) {
  constexpr int NP = (LB + LE - 1) / LE;
  static_for<0, NP>([&](auto PC) {
    constexpr int P = decltype(PC)::value;
    constexpr int HI = LB - P * LE;
    constexpr int LO = HI - LE > 0 ? HI - LE : 0;
    constexpr int WLO = LO;
    if constexpr (P == 0) {
      if constexpr (!PRELOADED) gmem_load<WLO, LE>(src, x, t);
    } else {
      lds_load<WLO, LE>(lds, x, t);
    }
    const uint32_t gbase = boff | ((t >> WLO) << (WLO + LE));
    fwd_layers<WLO, LE, HI, LO>(x, tw, nlog, gbase, p);
    if constexpr (P + 1 < NP) {
      constexpr int HI2 = LB - (P + 1) * LE;
      constexpr bool LOC = kWaveLocal<LB, LE, WLO, (HI2 - LE > 0 ? HI2 - LE : 0)>;
      xsync<LOC>(sync);
      lds_store<WLO, LE>(lds, x, t);
      xsync<LOC>(sync);
    }
  });
}

// Forward transforms of two blocks of one prime at once (x and y preloaded
// with gmem_load<FirstWindow<LB,LE>::value, LE>): per pass the layers of both
// share their twiddles; the exchanges go through LDS one block after the
// other.  On return x and y hold window [0, LE), as fwd_block leaves them.
// The LDS is taken to be free on entry (nothing of the workgroup still reads
// it: the kernel's or the work item's start) unless ENTRY_SYNC, in which
// case the first exchange begins with a barrier.
template <int LB, int LE, bool ENTRY_SYNC = false, class W, class TW, class S = WgSync>
__device__ __forceinline__ void fwd_block_dual(W (&x)[1 << LE], W (&y)[1 << LE], W *lds,
                                               const TW *__restrict__ tw, uint32_t nlog, uint32_t boff,
                                               typename Ident<W>::type p, uint32_t t, S &&sync = S()) {
  constexpr int NP = (LB + LE - 1) / LE;
  static_for<0, NP>([&](auto PC) {
    constexpr int P = decltype(PC)::value;
    constexpr int HI = LB - P * LE;
    constexpr int LO = HI - LE > 0 ? HI - LE : 0;
    constexpr int WLO = LO;
    const uint32_t gbase = boff | ((t >> WLO) << (WLO + LE));
    fwd_layers2<WLO, LE, HI, LO>(x, y, tw, nlog, gbase, p);
    if constexpr (P + 1 < NP) {
      constexpr int HI2 = LB - (P + 1) * LE;
      constexpr int WLO2 = HI2 - LE > 0 ? HI2 - LE : 0;
      constexpr bool LOC = kWaveLocal<LB, LE, WLO, WLO2>;
      if constexpr (P > 0 || ENTRY_SYNC) xsync<LOC>(sync);
      lds_store<WLO, LE>(lds, x, t);
      xsync<LOC>(sync);
      lds_load<WLO2, LE>(lds, x, t);
      xsync<LOC>(sync);
      lds_store<WLO, LE>(lds, y, t);
      xsync<LOC>(sync);
      lds_load<WLO2, LE>(lds, y, t);
    }
  });
}

// Inverse transform of index bits [0, LB) of one block starting from words
// in registers (window [0, LE), i = e | t << LE, each < 2p), written to
// global `dst` (block base).  TOP selects what happens at bit LB-1:
//   TOP_GS      plain GS layer (block of a larger residue), outputs in [0, 2p);
//   TOP_MERGED  the residue's top bit: merged N^-1 layer, canonical outputs;
//   TOP_SKIP    the residue's top bit left undone (negacylic_backward_lazy,
//               lib/ntt.cpp:192), outputs in [0, 2p).
enum { TOP_GS = 0, TOP_MERGED = 1, TOP_SKIP = 2 };

// Pass Q of the inverse of a 2^LB block, layers [lo, hi) in registers over
// window [wlo, wlo + LE): the last pass runs the top LE layers
// [LB - LE, LB); the ones before run [0, LB - LE) upwards in chunks of LE,
// the last chunk short, its window ending at its top layer (so that the
// exchanges up to the last pass stay inside each wave: kWaveLocal).  The
// layers run in the same order as one pass per layer.
template <int LB, int LE, int Q>
struct InvPass {
  static constexpr int np = (LB + LE - 1) / LE;
  static constexpr int top = LB - LE > 0 ? LB - LE : 0;
  static constexpr bool last = Q + 1 == np;
  static constexpr int lo = last ? top : Q * LE;
  static constexpr int hi = last ? LB : (Q * LE + LE < top ? Q * LE + LE : top);
  static constexpr int wlo = last ? top : (hi - LE > 0 ? hi - LE : 0);
};

// The first pass after pass 0 whose window starts at bit 6 or above: its
// threads' twiddles are wave-uniform (scalar loads), so from there to the
// end the inverse issues no vector load of its own.  inv_block runs its
// caller's hook there, before the pass's exchange load: vector loads the hook
// issues (the next residue's operand) are not waited for by the inverse.
template <int LB, int LE>
struct InvHookPass {
  static constexpr int find() {
    const int np = (LB + LE - 1) / LE, top = LB - LE > 0 ? LB - LE : 0;
    for (int q = 1; q < np; ++q) {  // InvPass<LB, LE, q>::wlo
      const int hi = q * LE + LE < top ? q * LE + LE : top;
      const int wlo = q + 1 == np ? top : (hi - LE > 0 ? hi - LE : 0);
      if (wlo >= 6) return q;
    }
    return np - 1;
  }
  static constexpr int value = find();
};

struct NoHook {
  __device__ __forceinline__ void operator()() const {}
};

template <int LB, int LE, int TOP, class W, class TW, class D, class S = WgSync, class H = NoHook>
__device__ __forceinline__ void inv_block(W (&x)[1 << LE], typename Ident<W>::type *__restrict__ dst,
                                          W *lds, const TW *__restrict__ tw, uint32_t nlog, uint32_t boff,
                                          const D &d, uint32_t t, S &&sync = S(), H &&hook = H()) {
  constexpr int NP = (LB + LE - 1) / LE;
  static_for<0, NP>([&](auto QC) {
    constexpr int Q = decltype(QC)::value;
    constexpr bool LASTQ = (Q + 1 == NP);
    constexpr int LO = InvPass<LB, LE, Q>::lo;
    constexpr int HI = InvPass<LB, LE, Q>::hi;
    constexpr int WLO = InvPass<LB, LE, Q>::wlo;
    if constexpr (Q == InvHookPass<LB, LE>::value) hook();
    if constexpr (Q > 0) lds_load<WLO, LE>(lds, x, t);
    const uint32_t gbase = boff | ((t >> WLO) << (WLO + LE));
    if constexpr (LASTQ && TOP == TOP_MERGED) {
      inv_layers<WLO, LE, HI - 1, LO>(x, tw, nlog, gbase, d.p);
      inv_last_layer<LE, LB - 1 - WLO>(x, d);
    } else if constexpr (LASTQ && TOP == TOP_SKIP) {
      inv_layers<WLO, LE, HI - 1, LO>(x, tw, nlog, gbase, d.p);
    } else {
      inv_layers<WLO, LE, HI, LO>(x, tw, nlog, gbase, d.p);
    }
    if constexpr (!LASTQ) {
      constexpr bool LOC = kWaveLocal<LB, LE, WLO, InvPass<LB, LE, Q + 1>::wlo>;
      xsync<LOC>(sync);
      lds_store<WLO, LE>(lds, x, t);
      xsync<LOC>(sync);
    } else if (dst != nullptr) {  // null: a sub-group with no residue of its own
      gmem_store<WLO, LE>(dst, x, t);
    }
  });
}

// The last forward pass leaves thread t with the 2^LE consecutive words
// t << LE | e (window [0, LE)), so a store instruction would touch one cache
// line per lane.  store_coalesced moves them through LDS to the first-window
// layout (thread t: words t | e << W0, consecutive across the lanes) and
// stores from there; load_coalesced is the converse for the inverse's first
// pass.  Measured on MI355X: a paired 2^14 forward 0.139 -> 0.122 ms per 1024
// residues, 2^13 0.066 -> 0.054 ms.  store_coalesced begins with a barrier (the
// LDS may still be read by the pass before); load_coalesced needs the LDS
// free and the caller's next exchange to begin with a barrier (inv_block's do).
template <int LB, int LE, class W>
__device__ __forceinline__ void store_coalesced(typename Ident<W>::type *__restrict__ dst, W (&x)[1 << LE],
                                                W *lds, uint32_t t) {
  constexpr int W0 = FirstWindow<LB, LE>::value;
  if constexpr (W0 == 0) {
    gmem_store<0, LE>(dst, x, t);
  } else {
    const uint32_t tt = phase_v(t);
    __syncthreads();
    lds_store<0, LE>(lds, x, tt);
    __syncthreads();
    lds_load<W0, LE>(lds, x, tt);
    gmem_store<W0, LE>(dst, x, tt);
  }
}

template <int LB, int LE, class W>
__device__ __forceinline__ void load_coalesced(const typename Ident<W>::type *__restrict__ src, W (&x)[1 << LE],
                                               W *lds, uint32_t t) {
  constexpr int W0 = FirstWindow<LB, LE>::value;
  if constexpr (W0 == 0) {
    gmem_load<0, LE>(src, x, t);
  } else {
    gmem_load<W0, LE>(src, x, t);
    lds_store<W0, LE>(lds, x, t);
    __syncthreads();
    lds_load<0, LE>(lds, x, t);
  }
}

// Runtime (lb, le) -> template instance.  Instantiated shapes: le = 3 for
// lb = 3, le = 4 for lb in [4, 14], le = 3 as an alternative for lb in
// [10, 13] and le = 5 (radix-32 passes, 512-thread workgroups at 2^14) for lb
// in [12, 14]; any other request falls back to le = 4 (clamp_le).
#define YA_LB_SWITCH(lb, le, CALL)                         \
  switch ((lb) * 8 + (le)) {                               \
    case 3 * 8 + 3: return CALL(3, 3);                     \
    case 4 * 8 + 4: return CALL(4, 4);                     \
    case 5 * 8 + 4: return CALL(5, 4);                     \
    case 6 * 8 + 4: return CALL(6, 4);                     \
    case 7 * 8 + 4: return CALL(7, 4);                     \
    case 8 * 8 + 4: return CALL(8, 4);                     \
    case 9 * 8 + 4: return CALL(9, 4);                     \
    case 10 * 8 + 3: return CALL(10, 3);                   \
    case 10 * 8 + 4: return CALL(10, 4);                   \
    case 11 * 8 + 3: return CALL(11, 3);                   \
    case 11 * 8 + 4: return CALL(11, 4);                   \
    case 12 * 8 + 3: return CALL(12, 3);                   \
    case 12 * 8 + 4: return CALL(12, 4);                   \
    case 12 * 8 + 5: return CALL(12, 5);                   \
    case 13 * 8 + 3: return CALL(13, 3);                   \
    case 13 * 8 + 4: return CALL(13, 4);                   \
    case 13 * 8 + 5: return CALL(13, 5);                   \
    case 14 * 8 + 4: return CALL(14, 4);                   \
    case 14 * 8 + 5: return CALL(14, 5);                   \
    default: return hipErrorInvalidValue;                  \
  }

inline int clamp_le(int lb, int le) {
  if (lb <= 3) return 3;
  if (le == 3 && lb >= 10 && lb <= 13) return 3;
  if (le == 5 && lb >= 12) return 5;
  return 4;
}

}  // namespace ya
