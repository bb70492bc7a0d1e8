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
  gmem_store<0, LE>(base, x, t);
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
  gmem_store<0, LE>(px, x, t);
  gmem_store<0, LE>(py, y, t);
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
  gmem_load<0, LE>(base, x, t);
  inv_block<LB, LE, TOP>(x, base, lds, d.inv, nlog, blk << LB, d, t);
}

// Column kernels for N = 2^(14+K): one thread per column c in [0, 2^14)
// holding the 2^K words c + e * 2^14.
constexpr int kTopThreads = 256;

template <int K>
__global__ void __launch_bounds__(kTopThreads)
k_fwd_top(uint64_t *__restrict__ data, const ModDesc *__restrict__ descs, uint32_t nmod) {
  constexpr int NLOG = kMaxBlockLog + K;
  const uint32_t cols_per_res_log = kMaxBlockLog - 8;  // 2^14 / 256 workgroups per residue
  const uint32_t res = blockIdx.x >> cols_per_res_log;
  const uint32_t t = ((blockIdx.x & ((1u << cols_per_res_log) - 1u)) << 8) | threadIdx.x;
  const ModDesc d = descs[res % nmod];
  uint64_t *base = data + ((size_t)res << NLOG);
  uint64_t x[1 << K];
  gmem_load<kMaxBlockLog, K>(base, x, t);
  fwd_layers<kMaxBlockLog, K, NLOG, kMaxBlockLog>(x, d.fwd, NLOG, 0u, d.p);
  gmem_store<kMaxBlockLog, K>(base, x, t);
}

template <int K, bool MERGED>
__global__ void __launch_bounds__(kTopThreads)
k_inv_top(uint64_t *__restrict__ data, const ModDesc *__restrict__ descs, uint32_t nmod) {
  constexpr int NLOG = kMaxBlockLog + K;
  const uint32_t cols_per_res_log = kMaxBlockLog - 8;
  const uint32_t res = blockIdx.x >> cols_per_res_log;
  const uint32_t t = ((blockIdx.x & ((1u << cols_per_res_log) - 1u)) << 8) | threadIdx.x;
  const ModDesc d = descs[res % nmod];
  uint64_t *base = data + ((size_t)res << NLOG);
  uint64_t x[1 << K];
  gmem_load<kMaxBlockLog, K>(base, x, t);
  inv_layers<kMaxBlockLog, K, NLOG - 1, kMaxBlockLog>(x, d.inv, NLOG, 0u, d.p);
  if constexpr (MERGED) inv_last_layer<K, K - 1>(x, d);
  gmem_store<kMaxBlockLog, K>(base, x, t);
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
  if (logn > kMaxBlockLog) {
    const int k = logn - kMaxBlockLog;
    const dim3 grid((unsigned)(nres << (kMaxBlockLog - 8)));
    switch (k) {
      case 1: hipLaunchKernelGGL(k_fwd_top<1>, grid, dim3(kTopThreads), 0, st, data, descs, (uint32_t)nmod); break;
      case 2: hipLaunchKernelGGL(k_fwd_top<2>, grid, dim3(kTopThreads), 0, st, data, descs, (uint32_t)nmod); break;
      case 3: hipLaunchKernelGGL(k_fwd_top<3>, grid, dim3(kTopThreads), 0, st, data, descs, (uint32_t)nmod); break;
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int lb = logn < kMaxBlockLog ? logn : kMaxBlockLog;
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
    const int k = logn - kMaxBlockLog;
    const dim3 grid((unsigned)((2 * npairs * nmod) << (kMaxBlockLog - 8)));
    switch (k) {
      case 1: hipLaunchKernelGGL(k_fwd_top<1>, grid, dim3(kTopThreads), 0, st, data, descs, (uint32_t)nmod); break;
      case 2: hipLaunchKernelGGL(k_fwd_top<2>, grid, dim3(kTopThreads), 0, st, data, descs, (uint32_t)nmod); break;
      case 3: hipLaunchKernelGGL(k_fwd_top<3>, grid, dim3(kTopThreads), 0, st, data, descs, (uint32_t)nmod); break;
      default: return hipErrorInvalidValue;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
#define YA_FWD_PAIR_CALL(LB, LE) fwd_pair_lb<LB, LE>(data, descs, logn, npairs, nmod, lazy, st)
  YA_LB_SWITCH(lb, le, YA_FWD_PAIR_CALL)
#undef YA_FWD_PAIR_CALL
}

static hipError_t backward_impl(uint64_t *data, const ModDesc *descs, int logn, size_t nres,
                                size_t nmod, bool lazy, NttShape shape, hipStream_t st) {
  if (logn < kMinLogN || logn > kMaxLogN || nres == 0 || nmod == 0) return hipErrorInvalidValue;
  const int lb = logn < kMaxBlockLog ? logn : kMaxBlockLog;
  const int le = clamp_le(lb, shape.le);
  hipError_t e;
  {
#define YA_INV_CALL(LB, LE) inv_lb<LB, LE>(data, descs, logn, nres, nmod, lazy, st)
    e = [&]() -> hipError_t { YA_LB_SWITCH(lb, le, YA_INV_CALL) }();
#undef YA_INV_CALL
  }
  if (e != hipSuccess || logn <= kMaxBlockLog) return e;
  const int k = logn - kMaxBlockLog;
  const dim3 grid((unsigned)(nres << (kMaxBlockLog - 8)));
  const dim3 blk(kTopThreads);
  switch (k * 2 + (lazy ? 0 : 1)) {
    case 2: hipLaunchKernelGGL((k_inv_top<1, false>), grid, blk, 0, st, data, descs, (uint32_t)nmod); break;
    case 3: hipLaunchKernelGGL((k_inv_top<1, true>), grid, blk, 0, st, data, descs, (uint32_t)nmod); break;
    case 4: hipLaunchKernelGGL((k_inv_top<2, false>), grid, blk, 0, st, data, descs, (uint32_t)nmod); break;
    case 5: hipLaunchKernelGGL((k_inv_top<2, true>), grid, blk, 0, st, data, descs, (uint32_t)nmod); break;
    case 6: hipLaunchKernelGGL((k_inv_top<3, false>), grid, blk, 0, st, data, descs, (uint32_t)nmod); break;
    case 7: hipLaunchKernelGGL((k_inv_top<3, true>), grid, blk, 0, st, data, descs, (uint32_t)nmod); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_backward(uint64_t *data, const ModDesc *descs, int logn, size_t nres, size_t nmod,
                           NttShape shape, hipStream_t st) {
  return backward_impl(data, descs, logn, nres, nmod, false, shape, st);
}

hipError_t launch_inverse_lazy(uint64_t *data, const ModDesc *descs, int logn, size_t nres,
                               hipStream_t st) {
  return backward_impl(data, descs, logn, nres, nres, true, NttShape{}, st);
}

}  // namespace ya
