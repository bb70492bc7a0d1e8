// ya_kernels_split.hip -- poly-multiplication of few residues, spread over
// many workgroups.
//
// The fused kernel gives each residue one workgroup; a launch of only a few
// residues (one request of 2^14 words x 8 primes is 8) then occupies 8 of the 256 CUs and its
// latency is that of one CU transforming 2^14 words three times.  Here a
// residue of N = 2^(S+K) words is split at index bit S:
//   1. column kernel: forward layers [S, S+K) of a (into c) and of b (into a
//      scratch residue, or in place in b when the caller owns b), one thread
//      per column c < 2^S holding the 2^K words c + e * 2^S;
//   2. block kernel: per block of 2^S words, forward layers [0, S) of both,
//      the Barrett product, inverse layers [0, S) (plain GS) into c;
//   3. column kernel: inverse layers [S, S+K) of c, the top one merged with
//      the N^-1 scaling.
// Every butterfly is the one of the single-kernel path with the same inputs,
// so the words are identical; there are 2^K times more workgroups.
#include <algorithm>

#include "ya_batch.hpp"
#include "ya_blocks.hpp"

namespace ya {

__device__ __forceinline__ void split_res(const SplitArgs &s, uint32_t v, uint32_t nlog,
                                          const uint64_t *&a, const uint64_t *&b, uint64_t *&c,
                                          uint64_t *&t, ModDesc &d) {
  const size_t off = (size_t)v << nlog;
  if (s.tab) {
    const BatchRes r = s.tab[v];
    a = r.a;
    b = r.b;
    c = r.c;
    d = load_desc(r.d);
  } else {
    a = s.a + off;
    b = s.b + off;
    c = s.c + off;
    d = s.descs[v % s.nmod];
  }
  t = s.t ? s.t + off : const_cast<uint64_t *>(b);
}

constexpr int kSplitColThreads = 256;

template <int S, int K>
__global__ void __launch_bounds__(kSplitColThreads)
k_split_fwd_cols(SplitArgs s) {
  constexpr int NLOG = S + K;
  constexpr int CW = S - 8;  // log2 workgroups per residue and operand
  const uint32_t blk = blockIdx.x;
  const uint32_t col = ((blk & ((1u << CW) - 1u)) << 8) | threadIdx.x;
  const uint32_t op = (blk >> CW) & 1u, v = blk >> (CW + 1);
  const uint64_t *a, *b;
  uint64_t *c, *t;
  ModDesc d;
  split_res(s, v, NLOG, a, b, c, t, d);
  uint64_t x[1 << K];
  gmem_load<S, K>(op ? b : a, x, col);
  fwd_layers<S, K, NLOG, S>(x, d.fwd, NLOG, 0u, d.p);
  gmem_store<S, K>(op ? t : c, x, col);
}

template <int S, int K>
__global__ void __launch_bounds__(kSplitColThreads)
k_split_inv_cols(SplitArgs s) {
  constexpr int NLOG = S + K;
  constexpr int CW = S - 8;
  const uint32_t blk = blockIdx.x;
  const uint32_t col = ((blk & ((1u << CW) - 1u)) << 8) | threadIdx.x;
  const uint32_t v = blk >> CW;
  const uint64_t *a, *b;
  uint64_t *c, *t;
  ModDesc d;
  split_res(s, v, NLOG, a, b, c, t, d);
  uint64_t x[1 << K];
  gmem_load<S, K>(c, x, col);
  inv_layers<S, K, NLOG - 1, S>(x, d.inv, NLOG, 0u, d.p);
  inv_last_layer<K, K - 1>(x, d);
  gmem_store<S, K>(c, x, col);
}

// Block j of 2^S words of residue v: c_j holds layers [S, S+K) of forward(a),
// t_j those of forward(b).
// (Four waves per SIMD: 512-thread workgroups of 2^13 blocks share a CU in
// pairs.)
template <int S, int LE>
__global__ void __launch_bounds__(1 << (S - LE), 4)
k_split_blocks(SplitArgs s, uint32_t nlog) {
  __shared__ uint64_t lds[LdsWords<S>::value];
  const uint32_t kb = nlog - S;
  const uint32_t j = blockIdx.x & ((1u << kb) - 1u), v = blockIdx.x >> kb;
  const uint64_t *a, *b;
  uint64_t *c, *tb;
  ModDesc d;
  split_res(s, v, nlog, a, b, c, tb, d);
  const uint32_t boff = j << S;
  uint64_t *cj = c + boff;
  const uint64_t *tj = tb + boff;
  const uint32_t t = threadIdx.x;
  uint64_t x[1 << LE], y[1 << LE];
  constexpr int W0 = FirstWindow<S, LE>::value;
  gmem_load<W0, LE>(cj, x, t);
  gmem_load<W0, LE>(tj, y, t);
  // the remaining forward layers of both blocks at once (shared twiddles)
  fwd_block_dual<S, LE>(x, y, lds, d.fwd, nlog, boff, d.p, t);
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(x[e], y[e], d.p, d.pn);
  const uint32_t t3 = phase_v(t);
  inv_block<S, LE, TOP_GS>(x, cj, lds, d.inv, nlog, boff, d, t3);
}

namespace {

template <int S, int K>
hipError_t split_launch(const SplitArgs &s, size_t nres, hipStream_t st) {
  constexpr int LE = S <= 9 ? 3 : 4;
  const dim3 cols((unsigned)(nres << (S - 8)));
  hipLaunchKernelGGL((k_split_fwd_cols<S, K>), dim3((unsigned)(nres << (S - 7))), dim3(kSplitColThreads), 0,
                     st, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_split_blocks<S, LE>), dim3((unsigned)(nres << K)), dim3(1u << (S - LE)), 0, st, s,
                     (uint32_t)(S + K));
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_split_inv_cols<S, K>), cols, dim3(kSplitColThreads), 0, st, s);
  return hipGetLastError();
}

}  // namespace

int split_choose_k(int logn, size_t nres, int cus) {
  if (logn < kSplitMinLog || logn > kSplitMaxLog || nres == 0) return 0;
  const int kmax = std::min(kSplitMaxK, logn - kSplitMinS);
  int k = kSplitMinK;
  while (k < kmax && (nres << k) < (size_t)cus) ++k;
  return k;
}

// 2^12 blocks at 2^15 and 2^16 (twice the workgroups of 2^13 blocks: 2^15 x 24
// primes 0.0535 -> 0.0428 ms, 2^16 x 24 0.0621 -> 0.0606, 2^16 x 4 0.0485 ->
// 0.0382; 2^11 blocks were slower again), 2^13 blocks up to 2^18 (at 2^17 x 12
// 2^12 and 2^13 blocks measured the same), 2^14 blocks at 2^19 (the column
// kernels take at most 5 index bits).
int large_ring_split_k(int logn) {
  if (logn <= kMaxBlockLog || logn > kMaxLogN) return 0;
  return logn <= 16 ? logn - 12 : logn <= 18 ? logn - 13 : logn - 14;
}

hipError_t launch_polymul_split(const SplitArgs &s, int logn, size_t nres, int k, hipStream_t st) {
  if (nres == 0 || nres > ((size_t)1 << 20)) return hipErrorInvalidValue;
  const int S = logn - k;
#define YA_SPLIT(SS, KK) \
  case SS * 16 + KK: return split_launch<SS, KK>(s, nres, st);
  switch (S * 16 + k) {
    YA_SPLIT(9, 2) YA_SPLIT(9, 3) YA_SPLIT(9, 4) YA_SPLIT(9, 5)
    YA_SPLIT(10, 2) YA_SPLIT(10, 3) YA_SPLIT(10, 4)
    YA_SPLIT(11, 2) YA_SPLIT(11, 3)
    YA_SPLIT(12, 2)
    // rings above 2^14 (large_ring_split_k)
    YA_SPLIT(12, 3) YA_SPLIT(12, 4) YA_SPLIT(13, 4) YA_SPLIT(13, 5)
    YA_SPLIT(14, 5)
    default: return hipErrorInvalidValue;
  }
#undef YA_SPLIT
}

}  // namespace ya
