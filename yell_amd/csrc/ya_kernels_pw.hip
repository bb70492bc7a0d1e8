// ya_kernels_pw.hip -- element-wise RNS operations.
#include "ya_kernels.hpp"

namespace ya {

// ---------------------------------------------------------------------------
// Pointwise ops: two words per thread.  Block row y is residue r = y % nmod
// of polynomial y / nmod (polynomials back to back in out and a); b's
// polynomial starts at (y / nmod) * b_stride (0: one b shared by all).
constexpr int kPwThreads = 256;

template <int OP>
__global__ void __launch_bounds__(kPwThreads)
k_pointwise(uint64_t *__restrict__ out, const uint64_t *__restrict__ a,
            const uint64_t *__restrict__ b, const ModDesc *__restrict__ descs, uint32_t n,
            uint32_t nmod, uint64_t b_stride) {
  const uint32_t y = blockIdx.y, r = y % nmod;
  const uint32_t i = (blockIdx.x * kPwThreads + threadIdx.x) * 2u;
  if (i >= n) return;
  const uint64_t p = descs[r].p, pn = descs[r].pn;
  const size_t k = (size_t)y * n + i;
  const size_t kb = (size_t)(y / nmod) * b_stride + (size_t)r * n + i;
  uint64_t o[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint64_t x = a[k + j];
    if constexpr (OP == PW_MUL) o[j] = mulmod_barrett(x, b[kb + j], p, pn);
    else if constexpr (OP == PW_ADD) o[j] = addmod(x, b[kb + j], p);
    else if constexpr (OP == PW_SUB) o[j] = submod(x, b[kb + j], p);
    else if constexpr (OP == PW_NEGATE) o[j] = p - x;
    else if constexpr (OP == PW_MULADD) o[j] = muladd_barrett(out[k + j], x, b[kb + j], p, pn);
    else o[j] = correct4p(x, p, p << 1);
  }
  out[k] = o[0];
  out[k + 1] = o[1];
}

// acc += a_j * b_j for j = 0..np-1, one Barrett muladd per pair (the
// reference's add_product_of applied pair after pair), two words per thread.
__global__ void __launch_bounds__(kPwThreads)
k_mac(uint64_t *__restrict__ acc, MacPairs pairs, int np, const ModDesc *__restrict__ descs,
      uint32_t n) {
  const uint32_t r = blockIdx.y;
  const uint32_t i = (blockIdx.x * kPwThreads + threadIdx.x) * 2u;
  if (i >= n) return;
  const uint64_t p = descs[r].p, pn = descs[r].pn;
  const size_t k = (size_t)r * n + i;
  uint64_t o0 = acc[k], o1 = acc[k + 1];
  for (int j = 0; j < np; ++j) {
    const uint64_t *a = pairs.a[j] + k, *b = pairs.b[j] + k;
    o0 = muladd_barrett(o0, a[0], b[0], p, pn);
    o1 = muladd_barrett(o1, a[1], b[1], p, pn);
  }
  acc[k] = o0;
  acc[k + 1] = o1;
}

hipError_t launch_mac(uint64_t *acc, const MacPairs &pairs, int npairs, const ModDesc *descs,
                      size_t n, size_t nres, hipStream_t st) {
  if (n < 2 || (n & 1) || nres == 0 || nres > 65535 || npairs < 0 || npairs > kMacPairs)
    return hipErrorInvalidValue;
  if (npairs == 0) return hipSuccess;
  const dim3 grid((unsigned)((n / 2 + kPwThreads - 1) / kPwThreads), (unsigned)nres);
  hipLaunchKernelGGL(k_mac, grid, dim3(kPwThreads), 0, st, acc, pairs, npairs, descs, (uint32_t)n);
  return hipGetLastError();
}

hipError_t launch_pointwise_many(PwOp op, uint64_t *out, const uint64_t *a, const uint64_t *b,
                                 const ModDesc *descs, size_t n, size_t nmod, size_t count, size_t b_stride,
                                 hipStream_t st) {
  if (n < 2 || (n & 1) || nmod == 0 || count == 0 || nmod * count > 65535) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((n / 2 + kPwThreads - 1) / kPwThreads), (unsigned)(nmod * count));
  const dim3 block(kPwThreads);
  const uint32_t n32 = (uint32_t)n, m32 = (uint32_t)nmod;
  const uint64_t bs = b_stride;
#define YA_PW_CASE(OP) \
  case OP: hipLaunchKernelGGL(k_pointwise<OP>, grid, block, 0, st, out, a, b, descs, n32, m32, bs); break;
  switch (op) {
    YA_PW_CASE(PW_MUL)
    YA_PW_CASE(PW_ADD)
    YA_PW_CASE(PW_SUB)
    YA_PW_CASE(PW_NEGATE)
    YA_PW_CASE(PW_MULADD)
    YA_PW_CASE(PW_CORRECT)
    default: return hipErrorInvalidValue;
  }
#undef YA_PW_CASE
  return hipGetLastError();
}

// One 128 KiB chunk of descriptor blockIdx.y per workgroup; each thread
// moves 16-byte vectors, four loads in flight before their stores (reads
// over PCIe have microseconds of latency).
constexpr uint32_t kXferThreads = 256, kXferChunkLog = 14;
typedef uint64_t xfer_v2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(kXferThreads) k_xfer(const XferDesc *__restrict__ descs) {
  const XferDesc x = descs[blockIdx.y];
  const uint64_t w0 = (uint64_t)blockIdx.x << kXferChunkLog;
  if (w0 >= x.words) return;
  const uint64_t w1 = min(x.words, w0 + (1ull << kXferChunkLog));
  const xfer_v2 *src = (const xfer_v2 *)x.src;
  xfer_v2 *dst = (xfer_v2 *)x.dst;
  const uint64_t v0 = w0 >> 1, v1 = w1 >> 1;  // 16-byte vector indices
  constexpr uint32_t U = 4;
  uint64_t v = v0 + threadIdx.x;
  for (; v + (U - 1) * kXferThreads < v1; v += U * kXferThreads) {
    xfer_v2 t[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) t[u] = src[v + u * kXferThreads];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) dst[v + u * kXferThreads] = t[u];
  }
  for (; v < v1; v += kXferThreads) dst[v] = src[v];
}

hipError_t launch_xfer(const XferDesc *descs, size_t n, size_t max_words, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (n > 65535 || max_words == 0) return hipErrorInvalidValue;
  const size_t chunks = (max_words + (1ull << kXferChunkLog) - 1) >> kXferChunkLog;
  if (chunks > 0x7fffffffu) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_xfer, dim3((unsigned)chunks, (unsigned)n), dim3(kXferThreads), 0, st, descs);
  return hipGetLastError();
}

hipError_t launch_pointwise(PwOp op, uint64_t *out, const uint64_t *a, const uint64_t *b,
                            const ModDesc *descs, size_t n, size_t nres, hipStream_t st) {
  return launch_pointwise_many(op, out, a, b, descs, n, nres, 1, 0, st);
}

}  // namespace ya
