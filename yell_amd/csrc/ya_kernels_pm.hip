// ya_kernels_pm.hip -- fused RNS poly-multiplication kernel.
#include <atomic>

#include "ya_batch.hpp"
#include "ya_blocks.hpp"

namespace ya {

// Fused c = backward(forward(a) (*) forward(b)) of one residue of 2^LB words
// by one workgroup.  forward(a) is parked in c at the positions this thread
// later reads back itself, so no other buffer is needed.
template <int LB, int LE, bool DUAL = false>
__device__ __forceinline__ void polymul_residue(uint64_t *__restrict__ c, const uint64_t *__restrict__ a,
                                                const uint64_t *__restrict__ b, const ModDesc &d,
                                                uint64_t *lds, uint32_t t) {
  uint64_t x[1 << LE], y[1 << LE];
  if constexpr (DUAL) {
    // both forward transforms at once (shared twiddles, no parking in c)
    constexpr int W0 = FirstWindow<LB, LE>::value;
    gmem_load<W0, LE>(a, x, t);
    gmem_load<W0, LE>(b, y, t);
    fwd_block_dual<LB, LE>(x, y, lds, (const TwC *)d.fwd, LB, 0u, d.p, t);
  } else if constexpr (LE <= 4) {
    // both operands are requested at once (b waits in y during forward(a)),
    // so the second transform does not stall on HBM latency again
    constexpr int W0 = FirstWindow<LB, LE>::value;
    gmem_load<W0, LE>(a, x, t);
    gmem_load<W0, LE>(b, y, t);
    fwd_block<LB, LE, true>(x, nullptr, lds, d.fwd, LB, 0u, d.p, t);
    gmem_store<0, LE>(c, x, t);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) x[e] = y[e];
    const uint32_t t2 = phase_v(t);
    fwd_block<LB, LE, true>(x, nullptr, lds, d.fwd + phase_zero_s(), LB, 0u, d.p, t2);
    gmem_load<0, LE>(c, y, t2);
  } else {
    fwd_block<LB, LE>(x, a, lds, d.fwd, LB, 0u, d.p, t);
    gmem_store<0, LE>(c, x, t);
    const uint32_t t2 = phase_v(t);
    fwd_block<LB, LE>(x, b, lds, d.fwd + phase_zero_s(), LB, 0u, d.p, t2);
    gmem_load<0, LE>(c, y, t2);
  }
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(y[e], x[e], d.p, d.pn);
  const uint32_t t3 = phase_v(t);
  inv_block<LB, LE, TOP_MERGED>(x, c, lds, (const TwC *)d.inv, LB, 0u, d, t3);
}

// Residue-major polynomials: residue res of the launch uses descs[res % nmod];
// the XCD ranges cover the residues in prime-major order, so an XCD's L2
// holds the twiddles of few primes.
template <int LB, int LE, bool DUAL>
__global__ void __launch_bounds__(1 << (LB - LE))
k_polymul(uint64_t *__restrict__ c, const uint64_t *__restrict__ a,
          const uint64_t *__restrict__ b, const ModDesc *__restrict__ descs, uint32_t nmod, uint32_t stagger) {
  __shared__ uint64_t lds[LdsWords<LB>::value];
  stagger_start(stagger);
  const uint32_t g = gridDim.x;
  const uint32_t v = xcd_order(blockIdx.x, g);
  const uint32_t count = g / nmod;  // polynomials in the launch
  const uint32_t cm = v / count, poly = v % count;
  // the residue index is wave-uniform: keep it (and the offsets derived from
  // it) in scalar registers, not in a vector pair live across the kernel
  const uint32_t res = __builtin_amdgcn_readfirstlane(poly * nmod + cm);
  const ModDesc d = descs[cm];
  const size_t off = (size_t)res << LB;
  polymul_residue<LB, LE, DUAL>(uniform_ptr(c + off), uniform_ptr(a + off), uniform_ptr(b + off), d, lds,
                                threadIdx.x);
}

// Looping form of the fused 2^LB kernel for launches of many residues, one
// workgroup per CU: workgroup w takes launch indices w, w + G, w + 2G, ...
// (G = the grid, a multiple of 8, so all of them map into w's XCD range of
// the prime-major order), and the next residue's b operand is loaded into y, which
// the inverse leaves free, from the inverse's first pass with wave-uniform
// twiddles on (InvHookPass): those loads overlap the inverse's last passes
// and stores instead of following them.  The same words as k_polymul.
template <int LB, int LE>
__global__ void __launch_bounds__(1 << (LB - LE))
k_polymul_loop(uint64_t *__restrict__ c, const uint64_t *__restrict__ a,
               const uint64_t *__restrict__ b, const ModDesc *__restrict__ descs, uint32_t nmod,
               uint32_t nres, uint32_t stagger) {
  static_assert(InvHookPass<LB, LE>::value > 0, "the next operand's load must follow the product");
  __shared__ uint64_t lds[LdsWords<LB>::value];
  stagger_start(stagger);
  constexpr int W0 = FirstWindow<LB, LE>::value;
  const uint32_t g = gridDim.x, count = nres / nmod;
  // launch index v -> position u of the prime-major order: prime cm =
  // u / count, polynomial u % count, stored at residue (u % count) * nmod + cm
  // (wave-uniform)
  uint32_t v = blockIdx.x;
  uint32_t u = xcd_order(v, nres);
  uint32_t cm = __builtin_amdgcn_readfirstlane(u / count);
  uint32_t res = __builtin_amdgcn_readfirstlane((u % count) * nmod + cm);
  uint64_t x[1 << LE], y[1 << LE];
  uint32_t t = threadIdx.x;
  gmem_load<W0, LE>(uniform_ptr(a + ((size_t)res << LB)), x, t);
  gmem_load<W0, LE>(uniform_ptr(b + ((size_t)res << LB)), y, t);
  for (;;) {
    const ModDesc d = descs[cm];
    // the previous residue's inverse may still read the LDS in other waves
    fwd_block_dual<LB, LE, true>(x, y, lds, (const TwC *)d.fwd, LB, 0u, d.p, t);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(y[e], x[e], d.p, d.pn);
    const uint32_t vn = v + g;
    const bool more = vn < nres;
    uint32_t cmn = 0, resn = 0;
    if (more) {
      const uint32_t un = xcd_order(vn, nres);
      cmn = __builtin_amdgcn_readfirstlane(un / count);
      resn = __builtin_amdgcn_readfirstlane((un % count) * nmod + cmn);
    }
    const uint32_t t3 = phase_v(threadIdx.x);
    inv_block<LB, LE, TOP_MERGED>(x, uniform_ptr(c + ((size_t)res << LB)), lds, (const TwC *)d.inv, LB, 0u, d,
                                  t3, WgSync(), [&] {
                                    if (more) gmem_load<W0, LE>(uniform_ptr(b + ((size_t)resn << LB)), y,
                                                                phase_v(threadIdx.x));
                                  });
    if (!more) break;
    t = phase_v(threadIdx.x);
    gmem_load<W0, LE>(uniform_ptr(a + ((size_t)resn << LB)), x, t);
    v = vn;
    cm = cmn;
    res = resn;
  }
}

// The looping form over a residue table (a batch plan of one degree class
// over arbitrary job pointers): launch index v reads entry xcd_order(v) of
// the prime-ordered table; otherwise as k_polymul_loop.
template <int LB, int LE>
__global__ void __launch_bounds__(1 << (LB - LE))
k_polymul_tab_loop(const BatchRes *__restrict__ tab, uint32_t nres, uint32_t stagger) {
  static_assert(InvHookPass<LB, LE>::value > 0, "the next operand's load must follow the product");
  __shared__ uint64_t lds[LdsWords<LB>::value];
  stagger_start(stagger);
  constexpr int W0 = FirstWindow<LB, LE>::value;
  const uint32_t g = gridDim.x;
  uint32_t v = blockIdx.x;
  BatchRes r = tab[xcd_order(v, nres)];
  uint64_t x[1 << LE], y[1 << LE];
  uint32_t t = threadIdx.x;
  gmem_load<W0, LE>(uniform_ptr(r.a), x, t);
  gmem_load<W0, LE>(uniform_ptr(r.b), y, t);
  for (;;) {
    const ModDesc d = load_desc(r.d);
    fwd_block_dual<LB, LE, true>(x, y, lds, (const TwC *)d.fwd, LB, 0u, d.p, t);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(y[e], x[e], d.p, d.pn);
    const uint32_t vn = v + g;
    const bool more = vn < nres;
    BatchRes rn = r;
    if (more) rn = tab[xcd_order(vn, nres)];
    const uint32_t t3 = phase_v(threadIdx.x);
    inv_block<LB, LE, TOP_MERGED>(x, uniform_ptr(r.c), lds, (const TwC *)d.inv, LB, 0u, d, t3, WgSync(), [&] {
      if (more) gmem_load<W0, LE>(uniform_ptr(rn.b), y, phase_v(threadIdx.x));
    });
    if (!more) break;
    t = phase_v(threadIdx.x);
    gmem_load<W0, LE>(uniform_ptr(rn.a), x, t);
    v = vn;
    r = rn;
  }
}

// One residue per workgroup from a residue table (a batch plan of one degree
// class); the table is ordered by prime, and the XCD ranges follow it.
template <int LB, int LE, bool DUAL>
__global__ void __launch_bounds__(1 << (LB - LE))
k_polymul_tab(const BatchRes *__restrict__ res, uint32_t nres, uint32_t stagger) {
  __shared__ uint64_t lds[LdsWords<LB>::value];
  stagger_start(stagger);
  const uint32_t v = xcd_order(blockIdx.x, nres);
  const BatchRes r = res[v];
  const ModDesc d = load_desc(r.d);
  polymul_residue<LB, LE, DUAL>(r.c, r.a, r.b, d, lds, threadIdx.x);
}

// Poly-multiplication of residues of 4 words, one per thread: forward (the
// reference's layers on bits 1, 0, corrected), the Barrett product, and the
// reference's backward at that size (both GS layers, then the merged layer on
// bit 1 again; see k_ntt4), so the words are the reference's sequence
// a.forward(); b.forward(); c = a * b; c.backward() at N = 4.
constexpr int kPm4Threads = 64;

__global__ void __launch_bounds__(kPm4Threads)
k_polymul4(uint64_t *__restrict__ c, const uint64_t *__restrict__ a, const uint64_t *__restrict__ b,
           const ModDesc *__restrict__ descs, uint32_t nres, uint32_t nmod) {
  const uint32_t v = blockIdx.x * kPm4Threads + threadIdx.x;
  if (v >= nres) return;
  const ModDesc d = descs[v % nmod];
  const size_t off = (size_t)v << 2;
  uint64_t x[4], y[4];
  gmem_load<0, 2>(a + off, x, 0u);
  gmem_load<0, 2>(b + off, y, 0u);
  fwd_layers2<0, 2, 2, 0>(x, y, d.fwd, 2u, 0u, d.p);
  correct_all<2>(x, d.p);
  correct_all<2>(y, d.p);
#pragma unroll
  for (int e = 0; e < 4; ++e) x[e] = mulmod_barrett(x[e], y[e], d.p, d.pn);
  inv_layers<0, 2, 2, 0>(x, d.inv, 2u, 0u, d.p);
  inv_last_layer<2, 1>(x, d);
  gmem_store<0, 2>(c + off, x, 0u);
}

namespace {

// The dual-forward form is instantiated for the shapes where it can pay
// (blocks of 2^10 and more, radix 8 / 16 passes).
template <int LB, int LE>
constexpr bool kDualShape = LB >= 10 && LE <= 4;

// Compute units of the current device (cached per device).
int current_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = cache[dev].load(std::memory_order_relaxed);
  if (c <= 0) {
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
    cache[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

// Workgroups of `kernel` one CU holds at once.
template <class K>
int resident_per_cu(K kernel, unsigned threads) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, (int)threads, 0) != hipSuccess || n <= 0) n = 1;
  return n;
}

// stagger_start's word for a launch of nres workgroups, `per_cu` resident per
// CU: the whole first wave waits in steps.  Measured at 2^14: 16 waves of
// workgroups 5% faster, 4 waves 1% slower, so only a launch of at least
// kStaggerMinWaves waves is staggered.
constexpr size_t kStaggerMinWaves = 8;
uint32_t stagger_word(int steps, int per_cu, size_t nres) {
  if (steps <= 0) return 0;
  const size_t first = (size_t)current_cus() * (size_t)per_cu;
  if (nres < kStaggerMinWaves * first || first > 0xffffff) return 0;
  return (uint32_t)(first << 8) | (uint32_t)(steps > 255 ? 255 : steps);
}

// The looping kernel is instantiated for the one shape of the batched
// 2^14 step; it runs launches of at least two residues per workgroup.
template <int LB, int LE>
constexpr bool kLoopShape = LB == 14 && LE == 4;

template <int LB, int LE>
hipError_t polymul_lb(uint64_t *c, const uint64_t *a, const uint64_t *b, const ModDesc *descs,
                      size_t nres, size_t nmod, NttShape shape, hipStream_t st) {
  constexpr unsigned kThreads = 1u << (LB - LE);
  if constexpr (kLoopShape<LB, LE>) {
    if (shape.dual) {
      static const int per_cu = resident_per_cu(k_polymul_loop<LB, LE>, kThreads);
      const size_t grid = ((size_t)current_cus() * (size_t)per_cu) & ~(size_t)7;
      if (grid >= 8 && nres >= 2 * grid && nres <= 0x7fffffff) {
        hipLaunchKernelGGL((k_polymul_loop<LB, LE>), dim3((unsigned)grid), dim3(kThreads), 0, st, c, a, b, descs,
                           (uint32_t)nmod, (uint32_t)nres, stagger_word(shape.stagger, per_cu, nres));
        return hipGetLastError();
      }
    }
  }
  if constexpr (kDualShape<LB, LE>) {
    if (shape.dual) {
      static const int per_cu = resident_per_cu(k_polymul<LB, LE, true>, kThreads);
      hipLaunchKernelGGL((k_polymul<LB, LE, true>), dim3((unsigned)nres), dim3(kThreads), 0, st, c, a, b, descs,
                         (uint32_t)nmod, stagger_word(shape.stagger, per_cu, nres));
      return hipGetLastError();
    }
  }
  static const int per_cu = resident_per_cu(k_polymul<LB, LE, false>, kThreads);
  hipLaunchKernelGGL((k_polymul<LB, LE, false>), dim3((unsigned)nres), dim3(kThreads), 0, st, c, a, b, descs,
                     (uint32_t)nmod, stagger_word(shape.stagger, per_cu, nres));
  return hipGetLastError();
}

template <int LB, int LE>
hipError_t polymul_tab_lb(const BatchRes *res, size_t nres, NttShape shape, hipStream_t st) {
  constexpr unsigned kThreads = 1u << (LB - LE);
  if constexpr (kLoopShape<LB, LE>) {
    if (shape.dual) {
      static const int per_cu = resident_per_cu(k_polymul_tab_loop<LB, LE>, kThreads);
      const size_t grid = ((size_t)current_cus() * (size_t)per_cu) & ~(size_t)7;
      if (grid >= 8 && nres >= 2 * grid) {
        hipLaunchKernelGGL((k_polymul_tab_loop<LB, LE>), dim3((unsigned)grid), dim3(kThreads), 0, st, res,
                           (uint32_t)nres, stagger_word(shape.stagger, per_cu, nres));
        return hipGetLastError();
      }
    }
  }
  if constexpr (kDualShape<LB, LE>) {
    if (shape.dual) {
      static const int per_cu = resident_per_cu(k_polymul_tab<LB, LE, true>, kThreads);
      hipLaunchKernelGGL((k_polymul_tab<LB, LE, true>), dim3((unsigned)nres), dim3(kThreads), 0, st, res,
                         (uint32_t)nres, stagger_word(shape.stagger, per_cu, nres));
      return hipGetLastError();
    }
  }
  static const int per_cu = resident_per_cu(k_polymul_tab<LB, LE, false>, kThreads);
  hipLaunchKernelGGL((k_polymul_tab<LB, LE, false>), dim3((unsigned)nres), dim3(kThreads), 0, st, res,
                     (uint32_t)nres, stagger_word(shape.stagger, per_cu, nres));
  return hipGetLastError();
}

}  // namespace

hipError_t launch_polymul_table(const BatchRes *res, size_t nres, int logn, NttShape shape,
                                hipStream_t st) {
  if (logn < kMinLogN || logn > kMaxBlockLog || nres == 0 || nres > 0x7fffffff)
    return hipErrorInvalidValue;
  const int le = clamp_le(logn, shape.le);
#define YA_PMT_CALL(LB, LE) polymul_tab_lb<LB, LE>(res, nres, shape, st)
  YA_LB_SWITCH(logn, le, YA_PMT_CALL)
#undef YA_PMT_CALL
}

hipError_t launch_polymul(uint64_t *c, const uint64_t *a, const uint64_t *b, const ModDesc *descs,
                          int logn, size_t nres, size_t nmod, NttShape shape, hipStream_t st) {
  if (logn < kMinLogN || logn > kMaxLogN || nres == 0 || nmod == 0 || nres % nmod)
    return hipErrorInvalidValue;
  if (logn > kMaxBlockLog) return hipErrorNotSupported;  // composed by the caller
  if (logn == 2) {
    if (nres > 0xffffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_polymul4, dim3((unsigned)((nres + kPm4Threads - 1) / kPm4Threads)), dim3(kPm4Threads), 0,
                       st, c, a, b, descs, (uint32_t)nres, (uint32_t)nmod);
    return hipGetLastError();
  }
  const int le = clamp_le(logn, shape.le);
#define YA_PM_CALL(LB, LE) polymul_lb<LB, LE>(c, a, b, descs, nres, nmod, shape, st)
  YA_LB_SWITCH(logn, le, YA_PM_CALL)
#undef YA_PM_CALL
}

}  // namespace ya
