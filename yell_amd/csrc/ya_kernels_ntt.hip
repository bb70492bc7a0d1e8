// ya_kernels_ntt.hip -- standalone forward / inverse NTT kernels (block
// kernels and the column kernels of rings larger than 2^14).
#include "ya_blocks.hpp"

namespace ya {

// ---------------------------------------------------------------------------
template <int LB, int LE, bool LAZY>
__global__ void __launch_bounds__(1 << (LB - LE))
k_fwd_block(uint64_t *__restrict__ data, const ModDesc *__restrict__ descs, uint32_t nlog,
            uint32_t nmod) {
  __shared__ uint64_t lds[LdsWords<LB>::value];
  const uint32_t nblk_log = nlog - LB;
  const uint32_t res = blockIdx.x >> nblk_log;
  const uint32_t blk = blockIdx.x & ((1u << nblk_log) - 1u);
  const ModDesc d = descs[res % nmod];
  uint64_t *base = data + ((size_t)res << nlog) + ((size_t)blk << LB);
  const uint32_t t = threadIdx.x;
  uint64_t x[1 << LE];
  fwd_block<LB, LE>(x, base, lds, d.fwd, nlog, blk << LB, d.p, t);
  if constexpr (!LAZY) correct_all<LE>(x, d.p);
  store_coalesced<LB, LE>(base, x, lds, t);
}

// Block blk (2^LB words) of residue r of polynomials 2q and 2q+1 of a
// many-polynomial launch in one workgroup: same prime and same block, so the
// same twiddles, run as the dual pass sequence of the fused kernel, each
// twiddle loaded once for both, twice the residues in flight per LDS buffer.
// A residue of 2^nlog > 2^LB words has had its top layers done by k_fwd_top.
// The same words as k_fwd_block on each.  Prime-major order over the XCDs.
template <int LB, int LE, bool LAZY>
__global__ void __launch_bounds__(1 << (LB - LE))
k_fwd_block_pair(uint64_t *__restrict__ data, const ModDesc *__restrict__ descs, uint32_t nmod,
                 uint32_t npairs, uint32_t nlog) {
  __shared__ uint64_t lds[LdsWords<LB>::value];
  const uint32_t nblk_log = nlog - LB;
  const uint32_t v = xcd_order(blockIdx.x, gridDim.x);
  const uint32_t per_prime = npairs << nblk_log;
  const uint32_t r = v / per_prime, rest = v % per_prime;
  const uint32_t q = rest >> nblk_log, blk = rest & ((1u << nblk_log) - 1u);
  const ModDesc d = descs[r];
  uint64_t *px = data + ((size_t)(2 * q * nmod + r) << nlog) + ((size_t)blk << LB);
  uint64_t *py = px + ((size_t)nmod << nlog);
  const uint32_t t = threadIdx.x;
  uint64_t x[1 << LE], y[1 << LE];
  constexpr int W0 = FirstWindow<LB, LE>::value;
  gmem_load<W0, LE>(px, x, t);
  gmem_load<W0, LE>(py, y, t);
  fwd_block_dual<LB, LE>(x, y, lds, d.fwd, nlog, blk << LB, d.p, t);
  if constexpr (!LAZY) {
    correct_all<LE>(x, d.p);
    correct_all<LE>(y, d.p);
  }
  store_coalesced<LB, LE>(px, x, lds, t);
  store_coalesced<LB, LE>(py, y, lds, t);
}

// ---- Fused steps of the reference example's RLWE flow over many polynomials
// (each gives the words of the unfused sequence of launches).

// Encryption step (example.cpp:41-44, 46-48) on residue r of polynomials 2q
// and 2q+1: acc = acc + add (operator+=, include/yell/ops.hpp:25-41) as the
// words are loaded, forward_lazy as k_fwd_block_pair, then each transform
// goes back through LDS to the load layout (thread t: words t | e << W0,
// coalesced) for acc = u * key + acc (add_product_of,
// include/yell/details/poly.hpp:119-132) with coalesced reads of u and key,
// and a coalesced store; key is one polynomial shared by all.  acc is read
// before the first barrier, so add, u and key must not alias it.
template <int LB, int LE, bool ADD>
__global__ void __launch_bounds__(1 << (LB - LE))
k_fwd_pair_mac(uint64_t *__restrict__ acc, const uint64_t *__restrict__ add, const uint64_t *__restrict__ u,
               const uint64_t *__restrict__ key, const ModDesc *__restrict__ descs, uint32_t nmod,
               uint32_t npairs) {
  typedef const __attribute__((address_space(1))) uint64_t gw;
  __shared__ uint64_t lds[LdsWords<LB>::value];
  const uint32_t v = xcd_order(blockIdx.x, gridDim.x);
  const uint32_t r = v / npairs, q = v % npairs;
  const ModDesc d = descs[r];
  const size_t ox = (size_t)(2 * q * nmod + r) << LB, oy = ox + ((size_t)nmod << LB);
  const uint32_t t = threadIdx.x;
  uint64_t x[1 << LE], y[1 << LE];
  constexpr int W0 = FirstWindow<LB, LE>::value;
  const uint32_t b0 = win_base<W0, LE>(t);
  gmem_load<W0, LE>(acc + ox, x, t);
  gmem_load<W0, LE>(acc + oy, y, t);
  if constexpr (ADD) {
    gw *ax = (gw *)(add + ox), *ay = (gw *)(add + oy);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) {
      const uint32_t i = b0 | ((uint32_t)e << W0);
      x[e] = addmod(x[e], ax[i], d.p);
      y[e] = addmod(y[e], ay[i], d.p);
    }
  }
  fwd_block_dual<LB, LE>(x, y, lds, d.fwd, LB, 0u, d.p, t);
  gw *kr = (gw *)(key + ((size_t)r << LB));
  // x: back to the load layout through LDS, then the product with u and key
  // (thread indices through phase_v so that the reads of u and key stay in
  // their phase instead of being hoisted, live, above the exchanges)
  const uint32_t tx = phase_v(t);
  __syncthreads();
  lds_store<0, LE>(lds, x, tx);
  __syncthreads();
  lds_load<W0, LE>(lds, x, tx);
  {
    gw *ur = (gw *)(u + ox);
    const uint32_t bx = win_base<W0, LE>(tx);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) {
      const uint32_t i = bx | ((uint32_t)e << W0);
      x[e] = muladd_barrett(x[e], ur[i], kr[i], d.p, d.pn);
    }
  }
  gmem_store<W0, LE>(acc + ox, x, tx);
  const uint32_t ty = phase_v(t);
  __syncthreads();
  lds_store<0, LE>(lds, y, ty);
  __syncthreads();
  lds_load<W0, LE>(lds, y, ty);
  {
    gw *ur = (gw *)(u + oy);
    const uint32_t by = win_base<W0, LE>(ty);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) {
      const uint32_t i = by | ((uint32_t)e << W0);
      y[e] = muladd_barrett(y[e], ur[i], kr[i], d.p, d.pn);
    }
  }
  gmem_store<W0, LE>(acc + oy, y, ty);
}

// Decryption step (example.cpp:55-57): out = backward(acc + x * key), the
// product (add_product_of; key shared by all polynomials) formed with
// coalesced reads (thread t: words t | e << W0) and passed through LDS to the
// inverse's first layout; one residue per workgroup as k_inv_block.  Every
// word of acc is read before the first barrier and out is written after the
// last, so out may be acc.
template <int LB, int LE>
__global__ void __launch_bounds__(1 << (LB - LE))
k_mac_inv(uint64_t *out, const uint64_t *acc, const uint64_t *__restrict__ xin, const uint64_t *__restrict__ key,
          const ModDesc *__restrict__ descs, uint32_t nmod) {
  typedef const __attribute__((address_space(1))) uint64_t gw;
  __shared__ uint64_t lds[LdsWords<LB>::value];
  const uint32_t res = blockIdx.x, r = res % nmod;
  const ModDesc d = descs[r];
  const size_t o = (size_t)res << LB;
  const uint32_t t = threadIdx.x;
  constexpr int W0 = FirstWindow<LB, LE>::value;
  gw *ar = (gw *)(acc + o), *xr = (gw *)(xin + o), *kr = (gw *)(key + ((size_t)r << LB));
  const uint32_t b0 = win_base<W0, LE>(t);
  uint64_t x[1 << LE];
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) {
    const uint32_t i = b0 | ((uint32_t)e << W0);
    x[e] = muladd_barrett(ar[i], xr[i], kr[i], d.p, d.pn);
  }
  lds_store<W0, LE>(lds, x, t);
  __syncthreads();
  lds_load<0, LE>(lds, x, t);
  inv_block<LB, LE, TOP_MERGED>(x, out + o, lds, d.inv, LB, 0u, d, t);
}

template <int LB, int LE, int TOP>
__global__ void __launch_bounds__(1 << (LB - LE))
k_inv_block(uint64_t *__restrict__ data, const ModDesc *__restrict__ descs, uint32_t nlog,
            uint32_t nmod) {
  __shared__ uint64_t lds[LdsWords<LB>::value];
  const uint32_t nblk_log = nlog - LB;
  const uint32_t res = blockIdx.x >> nblk_log;
  const uint32_t blk = blockIdx.x & ((1u << nblk_log) - 1u);
  const ModDesc d = descs[res % nmod];
  uint64_t *base = data + ((size_t)res << nlog) + ((size_t)blk << LB);
  const uint32_t t = threadIdx.x;
  uint64_t x[1 << LE];
  load_coalesced<LB, LE>(base, x, lds, t);
  inv_block<LB, LE, TOP>(x, base, lds, d.inv, nlog, blk << LB, d, t);
}

// Residues of 4 words, one per thread, in the reference's exact layer order
// at that size: forward (lib/ntt.cpp:122-147; its main loop is empty) runs
// the layers on bits 1 and 0.  The reference's inverse at N = 4 runs both of
// its special-cased loops (:161-187, bits 0 and 1) and its main loop not at
// all, so negacylic_backward_lazy does not skip the top layer there, and
// ntt<N>::backward then applies the merged layer on bit 1 once more
// (include/yell/details/ntt.hpp:151-159): not the inverse transform, but the
// reference's words, reproduced.  MODE 0 forward_lazy, 1 forward, 2 backward,
// 3 backward_lazy.
constexpr int kSmallThreads = 64;

template <int MODE>
__global__ void __launch_bounds__(kSmallThreads)
k_ntt4(uint64_t *__restrict__ data, const ModDesc *__restrict__ descs, uint32_t nres, uint32_t nmod) {
  const uint32_t v = blockIdx.x * kSmallThreads + threadIdx.x;
  if (v >= nres) return;
  const ModDesc d = descs[v % nmod];
  uint64_t *base = data + ((size_t)v << 2);
  uint64_t x[4];
  gmem_load<0, 2>(base, x, 0u);
  if constexpr (MODE <= 1) {
    fwd_layers<0, 2, 2, 0>(x, d.fwd, 2u, 0u, d.p);
    if constexpr (MODE == 1) correct_all<2>(x, d.p);
  } else {
    inv_layers<0, 2, 2, 0>(x, d.inv, 2u, 0u, d.p);
    if constexpr (MODE == 2) inv_last_layer<2, 1>(x, d);
  }
  gmem_store<0, 2>(base, x, 0u);
}

static hipError_t launch_ntt4(uint64_t *data, const ModDesc *descs, size_t nres, size_t nmod, int mode,
                              hipStream_t st) {
  if (nres > 0xffffffffu) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((nres + kSmallThreads - 1) / kSmallThreads));
  const uint32_t n32 = (uint32_t)nres, m32 = (uint32_t)nmod;
  switch (mode) {
    case 0: hipLaunchKernelGGL(k_ntt4<0>, grid, dim3(kSmallThreads), 0, st, data, descs, n32, m32); break;
    case 1: hipLaunchKernelGGL(k_ntt4<1>, grid, dim3(kSmallThreads), 0, st, data, descs, n32, m32); break;
    case 2: hipLaunchKernelGGL(k_ntt4<2>, grid, dim3(kSmallThreads), 0, st, data, descs, n32, m32); break;
    default: hipLaunchKernelGGL(k_ntt4<3>, grid, dim3(kSmallThreads), 0, st, data, descs, n32, m32); break;
  }
  return hipGetLastError();
}

// Column kernels for N = 2^(B+K): one thread per column c in [0, 2^B) holding
// the 2^K words c + e * 2^B (the layers on index bits >= B; the blocks of
// 2^B words take the rest).  Workgroups of 2^TW threads: 256, or 64 for a
// launch of few columns (the inverse of 24 residues of 2^16 in 2^12 columns: 384
// workgroups of 256 threads, 1.5 per CU, or 1536 of 64, 6 per CU).
constexpr int kTopThreads = 256;
constexpr int kTopThreadsLog = 8, kTopThreadsSmallLog = 6;
// workgroups of 256 threads below which the column kernels use 64
constexpr size_t kTopSmallBelow = 512;

template <int B, int K, int TW = kTopThreadsLog>
__global__ void __launch_bounds__(1 << TW)
k_fwd_top(uint64_t *__restrict__ data, const ModDesc *__restrict__ descs, uint32_t nmod) {
  constexpr int NLOG = B + K;
  constexpr uint32_t cols_per_res_log = B - TW;  // 2^B / 2^TW workgroups per residue
  const uint32_t res = blockIdx.x >> cols_per_res_log;
  const uint32_t t = ((blockIdx.x & ((1u << cols_per_res_log) - 1u)) << TW) | threadIdx.x;
  const ModDesc d = descs[res % nmod];
  uint64_t *base = data + ((size_t)res << NLOG);
  uint64_t x[1 << K];
  gmem_load<B, K>(base, x, t);
  fwd_layers<B, K, NLOG, B>(x, d.fwd, NLOG, 0u, d.p);
  gmem_store<B, K>(base, x, t);
}

template <int B, int K, bool MERGED, int TW = kTopThreadsLog>
__global__ void __launch_bounds__(1 << TW)
k_inv_top(uint64_t *__restrict__ data, const ModDesc *__restrict__ descs, uint32_t nmod) {
  constexpr int NLOG = B + K;
  constexpr uint32_t cols_per_res_log = B - TW;
  const uint32_t res = blockIdx.x >> cols_per_res_log;
  const uint32_t t = ((blockIdx.x & ((1u << cols_per_res_log) - 1u)) << TW) | threadIdx.x;
  const ModDesc d = descs[res % nmod];
  uint64_t *base = data + ((size_t)res << NLOG);
  uint64_t x[1 << K];
  gmem_load<B, K>(base, x, t);
  inv_layers<B, K, NLOG - 1, B>(x, d.inv, NLOG, 0u, d.p);
  if constexpr (MERGED) inv_last_layer<K, K - 1>(x, d);
  gmem_store<B, K>(base, x, t);
}

// Block size of a ring above 2^kMaxBlockLog: 2^14 words, or 2^13 / 2^12 when
// the launch would otherwise have fewer workgroups than the device has CUs
// (a few residues: one inverse of 24 residues of 2^16), keeping the
// column pass to at most 4 layers (5 at 2^19: 32 words per thread).
constexpr size_t kFillWorkgroups = 256;
static int large_block_log(int logn, size_t nres) {
  int lb = kMaxBlockLog;
  while (lb > 12 && logn - lb < 4 && (nres << (logn - lb)) < kFillWorkgroups) --lb;
  return lb;
}

template <int B, int TW>
static hipError_t fwd_top_bt(uint64_t *data, const ModDesc *descs, int k, size_t nres, size_t nmod,
                             hipStream_t st) {
  const dim3 grid((unsigned)(nres << (B - TW))), block(1u << TW);
  switch (k) {
    case 1: hipLaunchKernelGGL((k_fwd_top<B, 1, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    case 2: hipLaunchKernelGGL((k_fwd_top<B, 2, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    case 3: hipLaunchKernelGGL((k_fwd_top<B, 3, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    case 4: hipLaunchKernelGGL((k_fwd_top<B, 4, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    case 5: hipLaunchKernelGGL((k_fwd_top<B, 5, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int B>
static hipError_t fwd_top_b(uint64_t *data, const ModDesc *descs, int k, size_t nres, size_t nmod,
                            hipStream_t st) {
  if ((nres << (B - kTopThreadsLog)) < kTopSmallBelow)
    return fwd_top_bt<B, kTopThreadsSmallLog>(data, descs, k, nres, nmod, st);
  return fwd_top_bt<B, kTopThreadsLog>(data, descs, k, nres, nmod, st);
}

// The top K = logn - lb forward layers of nres residues (columns of 2^lb).
static hipError_t launch_fwd_top(uint64_t *data, const ModDesc *descs, int logn, int lb, size_t nres, size_t nmod,
                                 hipStream_t st) {
  const int k = logn - lb;
  switch (lb) {
    case 12: return fwd_top_b<12>(data, descs, k, nres, nmod, st);
    case 13: return fwd_top_b<13>(data, descs, k, nres, nmod, st);
    case 14: return fwd_top_b<14>(data, descs, k, nres, nmod, st);
    default: return hipErrorInvalidValue;
  }
}

template <int B, bool MERGED, int TW>
static hipError_t inv_top_bt(uint64_t *data, const ModDesc *descs, int k, size_t nres, size_t nmod,
                             hipStream_t st) {
  const dim3 grid((unsigned)(nres << (B - TW))), block(1u << TW);
  switch (k) {
    case 1: hipLaunchKernelGGL((k_inv_top<B, 1, MERGED, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    case 2: hipLaunchKernelGGL((k_inv_top<B, 2, MERGED, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    case 3: hipLaunchKernelGGL((k_inv_top<B, 3, MERGED, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    case 4: hipLaunchKernelGGL((k_inv_top<B, 4, MERGED, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    case 5: hipLaunchKernelGGL((k_inv_top<B, 5, MERGED, TW>), grid, block, 0, st, data, descs, (uint32_t)nmod); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int B, bool MERGED>
static hipError_t inv_top_b(uint64_t *data, const ModDesc *descs, int k, size_t nres, size_t nmod,
                            hipStream_t st) {
  if ((nres << (B - kTopThreadsLog)) < kTopSmallBelow)
    return inv_top_bt<B, MERGED, kTopThreadsSmallLog>(data, descs, k, nres, nmod, st);
  return inv_top_bt<B, MERGED, kTopThreadsLog>(data, descs, k, nres, nmod, st);
}

// The top K = logn - lb inverse layers (the last merged with N^-1 unless lazy).
static hipError_t launch_inv_top(uint64_t *data, const ModDesc *descs, int logn, int lb, size_t nres, size_t nmod,
                                 bool lazy, hipStream_t st) {
  const int k = logn - lb;
  switch (lb * 2 + (lazy ? 0 : 1)) {
    case 24: return inv_top_b<12, false>(data, descs, k, nres, nmod, st);
    case 25: return inv_top_b<12, true>(data, descs, k, nres, nmod, st);
    case 26: return inv_top_b<13, false>(data, descs, k, nres, nmod, st);
    case 27: return inv_top_b<13, true>(data, descs, k, nres, nmod, st);
    case 28: return inv_top_b<14, false>(data, descs, k, nres, nmod, st);
    case 29: return inv_top_b<14, true>(data, descs, k, nres, nmod, st);
    default: return hipErrorInvalidValue;
  }
}

namespace {

template <int LB, int LE>
hipError_t fwd_lb(uint64_t *data, const ModDesc *descs, int logn, size_t nres, size_t nmod,
                  bool lazy, hipStream_t st) {
  const dim3 grid((unsigned)(nres << (logn - LB)));
  const dim3 block(1u << (LB - LE));
  if (lazy) {
    hipLaunchKernelGGL((k_fwd_block<LB, LE, true>), grid, block, 0, st, data, descs, (uint32_t)logn, (uint32_t)nmod);
  } else {
    hipLaunchKernelGGL((k_fwd_block<LB, LE, false>), grid, block, 0, st, data, descs, (uint32_t)logn, (uint32_t)nmod);
  }
  return hipGetLastError();
}

template <int LB, int LE>
hipError_t fwd_pair_lb(uint64_t *data, const ModDesc *descs, int logn, size_t npairs, size_t nmod, bool lazy,
                       hipStream_t st) {
  if constexpr (LB >= kMinPairLog && LE <= 4) {
    const dim3 grid((unsigned)((npairs * nmod) << (logn - LB)));
    const dim3 block(1u << (LB - LE));
    if (lazy) {
      hipLaunchKernelGGL((k_fwd_block_pair<LB, LE, true>), grid, block, 0, st, data, descs, (uint32_t)nmod,
                         (uint32_t)npairs, (uint32_t)logn);
    } else {
      hipLaunchKernelGGL((k_fwd_block_pair<LB, LE, false>), grid, block, 0, st, data, descs, (uint32_t)nmod,
                         (uint32_t)npairs, (uint32_t)logn);
    }
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

template <int LB, int LE>
hipError_t inv_lb(uint64_t *data, const ModDesc *descs, int logn, size_t nres, size_t nmod,
                  bool lazy, hipStream_t st) {
  const dim3 grid((unsigned)(nres << (logn - LB)));
  const dim3 block(1u << (LB - LE));
  if (logn != LB) {
    hipLaunchKernelGGL((k_inv_block<LB, LE, TOP_GS>), grid, block, 0, st, data, descs, (uint32_t)logn, (uint32_t)nmod);
  } else if (lazy) {
    hipLaunchKernelGGL((k_inv_block<LB, LE, TOP_SKIP>), grid, block, 0, st, data, descs, (uint32_t)logn, (uint32_t)nmod);
  } else {
    hipLaunchKernelGGL((k_inv_block<LB, LE, TOP_MERGED>), grid, block, 0, st, data, descs, (uint32_t)logn, (uint32_t)nmod);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_forward(uint64_t *data, const ModDesc *descs, int logn, size_t nres, size_t nmod,
                          bool lazy, NttShape shape, hipStream_t st) {
  if (logn < kMinLogN || logn > kMaxLogN || nres == 0 || nmod == 0) return hipErrorInvalidValue;
  if (logn == 2) return launch_ntt4(data, descs, nres, nmod, lazy ? 0 : 1, st);
  const int lb = logn <= kMaxBlockLog ? logn : large_block_log(logn, nres);
  if (logn > lb) {
    const hipError_t e = launch_fwd_top(data, descs, logn, lb, nres, nmod, st);
    if (e != hipSuccess) return e;
  }
  const int le = clamp_le(lb, shape.le);
#define YA_FWD_CALL(LB, LE) fwd_lb<LB, LE>(data, descs, logn, nres, nmod, lazy, st)
  YA_LB_SWITCH(lb, le, YA_FWD_CALL)
#undef YA_FWD_CALL
}

bool forward_pairs_ok(int logn, NttShape shape) {
  const int lb = logn < kMaxBlockLog ? logn : kMaxBlockLog;
  return logn >= kMinPairLog && logn <= kMaxLogN && clamp_le(lb, shape.le) <= 4;
}

hipError_t launch_forward_pairs(uint64_t *data, const ModDesc *descs, int logn, size_t npairs, size_t nmod,
                                bool lazy, NttShape shape, hipStream_t st) {
  if (logn < kMinPairLog || logn > kMaxLogN || npairs == 0 || nmod == 0) return hipErrorInvalidValue;
  if (npairs * nmod > 0x7fffffffu >> (logn - 3)) return hipErrorInvalidValue;  // grid in work-items
  const int lb = logn < kMaxBlockLog ? logn : kMaxBlockLog;
  const int le = clamp_le(lb, shape.le);
  if (le > 4) return hipErrorInvalidValue;
  if (logn > kMaxBlockLog) {
    // top layers of every residue of the 2 * npairs polynomials, then the
    // 2^kMaxBlockLog blocks in pairs
    const hipError_t e = launch_fwd_top(data, descs, logn, kMaxBlockLog, 2 * npairs * nmod, nmod, st);
    if (e != hipSuccess) return e;
  }
#define YA_FWD_PAIR_CALL(LB, LE) fwd_pair_lb<LB, LE>(data, descs, logn, npairs, nmod, lazy, st)
  YA_LB_SWITCH(lb, le, YA_FWD_PAIR_CALL)
#undef YA_FWD_PAIR_CALL
}

namespace {

template <int LB, int LE>
hipError_t fwd_pair_mac_lb(uint64_t *acc, const uint64_t *add, const uint64_t *u, const uint64_t *key,
                           const ModDesc *descs, size_t npairs, size_t nmod, hipStream_t st) {
  if constexpr (LB >= kMinPairLog && LE <= 4) {
    const dim3 grid((unsigned)(npairs * nmod));
    const dim3 block(1u << (LB - LE));
    if (add)
      hipLaunchKernelGGL((k_fwd_pair_mac<LB, LE, true>), grid, block, 0, st, acc, add, u, key, descs, (uint32_t)nmod,
                         (uint32_t)npairs);
    else
      hipLaunchKernelGGL((k_fwd_pair_mac<LB, LE, false>), grid, block, 0, st, acc, add, u, key, descs,
                         (uint32_t)nmod, (uint32_t)npairs);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

template <int LB, int LE>
hipError_t mac_inv_lb(uint64_t *out, const uint64_t *acc, const uint64_t *x, const uint64_t *key,
                      const ModDesc *descs, size_t nres, size_t nmod, hipStream_t st) {
  hipLaunchKernelGGL((k_mac_inv<LB, LE>), dim3((unsigned)nres), dim3(1u << (LB - LE)), 0, st, out, acc, x, key,
                     descs, (uint32_t)nmod);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_forward_pairs_mac(uint64_t *acc, const uint64_t *add, const uint64_t *u, const uint64_t *key,
                                    const ModDesc *descs, int logn, size_t npairs, size_t nmod, NttShape shape,
                                    hipStream_t st) {
  if (logn < kMinPairLog || logn > kMaxBlockLog || npairs == 0 || nmod == 0) return hipErrorInvalidValue;
  if (npairs * nmod > 0x7fffffffu >> (logn - 3)) return hipErrorInvalidValue;  // grid in work-items
  const int le = clamp_le(logn, shape.le);
  if (le > 4) return hipErrorInvalidValue;
#define YA_FWD_PAIR_MAC_CALL(LB, LE) fwd_pair_mac_lb<LB, LE>(acc, add, u, key, descs, npairs, nmod, st)
  YA_LB_SWITCH(logn, le, YA_FWD_PAIR_MAC_CALL)
#undef YA_FWD_PAIR_MAC_CALL
}

hipError_t launch_mac_inverse(uint64_t *out, const uint64_t *acc, const uint64_t *x, const uint64_t *key,
                              const ModDesc *descs, int logn, size_t nres, size_t nmod, NttShape shape,
                              hipStream_t st) {
  if (logn < 3 || logn > kMaxBlockLog || nres == 0 || nmod == 0 || nres > 0x7fffffffu)
    return hipErrorInvalidValue;
  const int le = clamp_le(logn, shape.le);
#define YA_MAC_INV_CALL(LB, LE) mac_inv_lb<LB, LE>(out, acc, x, key, descs, nres, nmod, st)
  YA_LB_SWITCH(logn, le, YA_MAC_INV_CALL)
#undef YA_MAC_INV_CALL
}

static hipError_t backward_impl(uint64_t *data, const ModDesc *descs, int logn, size_t nres,
                                size_t nmod, bool lazy, NttShape shape, hipStream_t st) {
  if (logn < kMinLogN || logn > kMaxLogN || nres == 0 || nmod == 0) return hipErrorInvalidValue;
  if (logn == 2) return launch_ntt4(data, descs, nres, nmod, lazy ? 3 : 2, st);
  const int lb = logn <= kMaxBlockLog ? logn : large_block_log(logn, nres);
  const int le = clamp_le(lb, shape.le);
  hipError_t e;
  {
#define YA_INV_CALL(LB, LE) inv_lb<LB, LE>(data, descs, logn, nres, nmod, lazy, st)
    e = [&]() -> hipError_t { YA_LB_SWITCH(lb, le, YA_INV_CALL) }();
#undef YA_INV_CALL
  }
  if (e != hipSuccess || logn <= lb) return e;
  return launch_inv_top(data, descs, logn, lb, nres, nmod, lazy, st);
}

hipError_t launch_backward(uint64_t *data, const ModDesc *descs, int logn, size_t nres, size_t nmod,
                           NttShape shape, hipStream_t st) {
  return backward_impl(data, descs, logn, nres, nmod, false, shape, st);
}

hipError_t launch_inverse_lazy(uint64_t *data, const ModDesc *descs, int logn, size_t nres,
                               hipStream_t st) {
  return backward_impl(data, descs, logn, nres, nres, true, NttShape{}, st);
}

// ---- The large-ring multiply-accumulate (launch_pair_products,
// launch_sum_backward).

// Top forward layers (index bits [B, B+K)) of polynomial blockIdx.y of the
// pair table (a_{y/2} or b_{y/2}), in place, one thread per column.
template <int B, int K>
__global__ void __launch_bounds__(kTopThreads)
k_fwd_top_pairs(PairPtrs pp, const ModDesc *__restrict__ descs) {
  constexpr int NLOG = B + K;
  constexpr uint32_t cols_per_res_log = B - 8;
  const uint32_t res = blockIdx.x >> cols_per_res_log;
  const uint32_t t = ((blockIdx.x & ((1u << cols_per_res_log) - 1u)) << 8) | threadIdx.x;
  const uint32_t q = blockIdx.y;
  const ModDesc d = descs[res];
  uint64_t *base = ((q & 1u) ? pp.b[q >> 1] : pp.a[q >> 1]) + ((size_t)res << NLOG);
  uint64_t x[1 << K];
  gmem_load<B, K>(base, x, t);
  fwd_layers<B, K, NLOG, B>(x, d.fwd, NLOG, 0u, d.p);
  gmem_store<B, K>(base, x, t);
}

// Block blk of residue r of pair j: the remaining forward layers of a_j and
// b_j as one dual pass sequence (shared twiddles), their Barrett product
// (canonical from the lazy words), stored to a_j in natural order.  Residue
// (prime) major over the XCDs, then pair, then block.
// (Four waves per SIMD: at 2^13, two 512-thread workgroups share a CU.)
template <int LB, int LE>
__global__ void __launch_bounds__(1 << (LB - LE), 4)
k_pair_product(PairPtrs pp, const ModDesc *__restrict__ descs, uint32_t npairs, uint32_t nlog) {
  __shared__ uint64_t lds[LdsWords<LB>::value];
  const uint32_t nblk_log = nlog - LB;
  const uint32_t v = xcd_order(blockIdx.x, gridDim.x);
  const uint32_t per_prime = npairs << nblk_log;
  const uint32_t r = v / per_prime, rest = v % per_prime;
  const uint32_t j = rest >> nblk_log, blk = rest & ((1u << nblk_log) - 1u);
  const ModDesc d = descs[r];
  const size_t off = ((size_t)r << nlog) + ((size_t)blk << LB);
  uint64_t *px = uniform_ptr(pp.a[j] + off);
  const uint64_t *py = uniform_ptr(pp.b[j] + off);
  const uint32_t t = threadIdx.x;
  uint64_t x[1 << LE], y[1 << LE];
  constexpr int W0 = FirstWindow<LB, LE>::value;
  gmem_load<W0, LE>(px, x, t);
  gmem_load<W0, LE>(py, y, t);
  fwd_block_dual<LB, LE>(x, y, lds, d.fwd, nlog, blk << LB, d.p, t);
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(x[e], y[e], d.p, d.pn);
  store_coalesced<LB, LE>(px, x, lds, t);
}

// Block blk of residue r: acc + sum_j a_j with coalesced reads (thread t:
// words t | e << W0), one LDS exchange into the inverse's first layout, the
// inverse layers of the block (TOP_MERGED when the block is the residue).
// acc is read before the first barrier and written after the last, by this
// workgroup only.  ACC false: acc is not read (the sum starts at a_0, the
// words of 0 + a_0 for a canonical a_0) and only written.
template <int LB, int LE, int TOP, bool ACC>
__global__ void __launch_bounds__(1 << (LB - LE))
k_sum_inv(uint64_t *__restrict__ acc, PairPtrs pp, uint32_t npairs, const ModDesc *__restrict__ descs,
          uint32_t nlog) {
  typedef const __attribute__((address_space(1))) uint64_t gw;
  __shared__ uint64_t lds[LdsWords<LB>::value];
  const uint32_t nblk_log = nlog - LB;
  const uint32_t r = blockIdx.x >> nblk_log, blk = blockIdx.x & ((1u << nblk_log) - 1u);
  const ModDesc d = descs[r];
  const size_t off = ((size_t)r << nlog) + ((size_t)blk << LB);
  const uint32_t t = threadIdx.x;
  constexpr int W0 = FirstWindow<LB, LE>::value;
  const uint32_t b0 = win_base<W0, LE>(t);
  uint64_t x[1 << LE];
  uint32_t j = 0;
  if constexpr (ACC) {
    gw *ar = (gw *)uniform_ptr(acc + off);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) x[e] = ar[b0 | ((uint32_t)e << W0)];
  } else {
    gw *a0 = (gw *)uniform_ptr(pp.a[0] + off);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) x[e] = a0[b0 | ((uint32_t)e << W0)];
    j = 1;
  }
  // U products' loads in flight at a time (4 with 256-thread workgroups, 2
  // within the 128 registers of the larger ones), then one at a time
  constexpr int U = LB <= 12 ? 4 : 2;
  for (; j + U <= npairs; j += U) {
    uint64_t v[U][1 << LE];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      gw *pu = (gw *)uniform_ptr(pp.a[j + u] + off);
#pragma unroll
      for (int e = 0; e < (1 << LE); ++e) v[u][e] = pu[b0 | ((uint32_t)e << W0)];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int e = 0; e < (1 << LE); ++e) x[e] = addmod(x[e], v[u][e], d.p);
  }
  for (; j < npairs; ++j) {
    gw *pj = (gw *)uniform_ptr(pp.a[j] + off);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) x[e] = addmod(x[e], pj[b0 | ((uint32_t)e << W0)], d.p);
  }
  lds_store<W0, LE>(lds, x, t);
  __syncthreads();
  lds_load<0, LE>(lds, x, t);
  inv_block<LB, LE, TOP>(x, acc + off, lds, d.inv, nlog, blk << LB, d, t);
}

template <int B>
static hipError_t fwd_top_pairs_b(const PairPtrs &pp, int npairs, const ModDesc *descs, int k, size_t nmod,
                                  hipStream_t st) {
  const dim3 grid((unsigned)(nmod << (B - 8)), (unsigned)(2 * npairs));
  switch (k) {
    case 1: hipLaunchKernelGGL((k_fwd_top_pairs<B, 1>), grid, dim3(kTopThreads), 0, st, pp, descs); break;
    case 2: hipLaunchKernelGGL((k_fwd_top_pairs<B, 2>), grid, dim3(kTopThreads), 0, st, pp, descs); break;
    case 3: hipLaunchKernelGGL((k_fwd_top_pairs<B, 3>), grid, dim3(kTopThreads), 0, st, pp, descs); break;
    case 4: hipLaunchKernelGGL((k_fwd_top_pairs<B, 4>), grid, dim3(kTopThreads), 0, st, pp, descs); break;
    case 5: hipLaunchKernelGGL((k_fwd_top_pairs<B, 5>), grid, dim3(kTopThreads), 0, st, pp, descs); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int LB>
static hipError_t pair_product_lb(const PairPtrs &pp, int npairs, const ModDesc *descs, int logn, size_t nmod,
                                  hipStream_t st) {
  const dim3 grid((unsigned)((nmod * npairs) << (logn - LB)));
  hipLaunchKernelGGL((k_pair_product<LB, 4>), grid, dim3(1u << (LB - 4)), 0, st, pp, descs, (uint32_t)npairs,
                     (uint32_t)logn);
  return hipGetLastError();
}

hipError_t launch_pair_products(const PairPtrs &pp, int npairs, const ModDesc *descs, int logn, size_t nmod,
                                int lb, hipStream_t st) {
  if (logn < kMinPairLog || logn > kMaxLogN || npairs <= 0 || npairs > kMacPairs || nmod == 0)
    return hipErrorInvalidValue;
  if (((nmod * npairs) << (logn - kMinPairLog)) > 0x7fffffffu) return hipErrorInvalidValue;
  if (logn <= lb) lb = logn;
  if ((lb != 12 && lb != 13 && lb != 14) || logn - lb > 5) return hipErrorInvalidValue;
  if (logn > lb) {
    const hipError_t e = lb == 12 ? fwd_top_pairs_b<12>(pp, npairs, descs, logn - lb, nmod, st)
                         : lb == 13 ? fwd_top_pairs_b<13>(pp, npairs, descs, logn - lb, nmod, st)
                                    : fwd_top_pairs_b<14>(pp, npairs, descs, logn - lb, nmod, st);
    if (e != hipSuccess) return e;
  }
  return lb == 12 ? pair_product_lb<12>(pp, npairs, descs, logn, nmod, st)
         : lb == 13 ? pair_product_lb<13>(pp, npairs, descs, logn, nmod, st)
                    : pair_product_lb<14>(pp, npairs, descs, logn, nmod, st);
}

template <int LB, int TOP, int LE = 4>
static hipError_t sum_inv_lb(uint64_t *acc, const PairPtrs &pp, int npairs, const ModDesc *descs, int logn,
                             size_t nmod, bool accumulate, hipStream_t st) {
  const dim3 grid((unsigned)(nmod << (logn - LB)));
  if (accumulate)
    hipLaunchKernelGGL((k_sum_inv<LB, LE, TOP, true>), grid, dim3(1u << (LB - LE)), 0, st, acc, pp,
                       (uint32_t)npairs, descs, (uint32_t)logn);
  else
    hipLaunchKernelGGL((k_sum_inv<LB, LE, TOP, false>), grid, dim3(1u << (LB - LE)), 0, st, acc, pp,
                       (uint32_t)npairs, descs, (uint32_t)logn);
  return hipGetLastError();
}

hipError_t launch_sum_backward(uint64_t *acc, const PairPtrs &pp, int npairs, const ModDesc *descs, int logn,
                               size_t nmod, bool accumulate, hipStream_t st) {
  if (logn < kMinPairLog || logn > kMaxLogN || npairs < 0 || npairs > kMacPairs || nmod == 0 ||
      (nmod << (logn - 12)) > 0x7fffffffu || (!accumulate && npairs < 1))
    return hipErrorInvalidValue;
  // blocks of 2^14, or 2^13 / 2^12 to give a launch of few residues more
  // workgroups (the inverse's column pass then takes the rest)
  const int lb = logn <= kMaxBlockLog ? logn : large_block_log(logn, nmod);
  hipError_t e;
  if (logn == lb) {
    e = lb == 13 ? sum_inv_lb<13, TOP_MERGED>(acc, pp, npairs, descs, logn, nmod, accumulate, st)
                 : sum_inv_lb<14, TOP_MERGED>(acc, pp, npairs, descs, logn, nmod, accumulate, st);
    return e;
  }
  // radix-8 passes below 2^14: twice the waves of radix-16 ones, for a pass
  // that is short of waves (the sum's loads and the inverse's dependency
  // chains need them)
  switch (lb) {
    case 12: e = sum_inv_lb<12, TOP_GS, 3>(acc, pp, npairs, descs, logn, nmod, accumulate, st); break;
    case 13: e = sum_inv_lb<13, TOP_GS, 3>(acc, pp, npairs, descs, logn, nmod, accumulate, st); break;
    default: e = sum_inv_lb<14, TOP_GS>(acc, pp, npairs, descs, logn, nmod, accumulate, st); break;
  }
  if (e != hipSuccess) return e;
  return launch_inv_top(acc, descs, logn, lb, nmod, nmod, false, st);
}

}  // namespace ya
