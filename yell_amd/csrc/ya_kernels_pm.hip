// ya_kernels_pm.hip -- fused RNS poly-multiplication kernel.
#include "ya_blocks.hpp"

namespace ya {

// Fused c = backward(forward(a) (*) forward(b)) for one residue of 2^LB words
// per workgroup.  forward(a) is parked in c at the positions this thread
// later reads back itself, so no other buffer is needed.
template <int LB, int LE>
__global__ void __launch_bounds__(1 << (LB - LE))
k_polymul(uint64_t *__restrict__ c, const uint64_t *__restrict__ a,
          const uint64_t *__restrict__ b, const ModDesc *__restrict__ descs, uint32_t nmod) {
  __shared__ uint64_t lds[LdsWords<LB>::value];
  // XCD-aware order: workgroups b, b+8, b+16, ... run on one XCD (the
  // hardware deals consecutive workgroups to consecutive XCDs); give each XCD
  // a contiguous range of the residues in prime-major order, so its L2 holds
  // the twiddles of few primes.
  const uint32_t g = gridDim.x, xcd = blockIdx.x % 8u, pos = blockIdx.x / 8u;
  const uint32_t v = xcd * (g / 8u) + min(xcd, g % 8u) + pos;
  const uint32_t count = g / nmod;             // polynomials in the launch
  const uint32_t cm = v / count, poly = v % count;
  const uint32_t res = poly * nmod + cm;
  const ModDesc d = descs[cm];
  const size_t off = (size_t)res << LB;
  const uint32_t t = threadIdx.x;
  uint64_t x[1 << LE], y[1 << LE];
  if constexpr (LE <= 4) {
    // both operands are requested at once (b waits in y during forward(a)),
    // so the second transform does not stall on HBM latency again
    constexpr int W0 = FirstWindow<LB, LE>::value;
    gmem_load<W0, LE>(a + off, x, t);
    gmem_load<W0, LE>(b + off, y, t);
    fwd_block<LB, LE, true>(x, nullptr, lds, d.fwd, LB, 0u, d.p, t);
    gmem_store<0, LE>(c + off, x, t);
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) x[e] = y[e];
    const uint32_t t2 = phase_v(t);
    fwd_block<LB, LE, true>(x, nullptr, lds, d.fwd + phase_zero_s(), LB, 0u, d.p, t2);
    gmem_load<0, LE>(c + off, y, t2);
  } else {
    fwd_block<LB, LE>(x, a + off, lds, d.fwd, LB, 0u, d.p, t);
    gmem_store<0, LE>(c + off, x, t);
    const uint32_t t2 = phase_v(t);
    fwd_block<LB, LE>(x, b + off, lds, d.fwd + phase_zero_s(), LB, 0u, d.p, t2);
    gmem_load<0, LE>(c + off, y, t2);
  }
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(y[e], x[e], d.p, d.pn);
  const uint32_t t3 = phase_v(t);
  inv_block<LB, LE, TOP_MERGED>(x, c + off, lds, d.inv, LB, 0u, d, t3);
}

namespace {

template <int LB, int LE>
hipError_t polymul_lb(uint64_t *c, const uint64_t *a, const uint64_t *b, const ModDesc *descs,
                      size_t nres, size_t nmod, hipStream_t st) {
  hipLaunchKernelGGL((k_polymul<LB, LE>), dim3((unsigned)nres), dim3(1u << (LB - LE)), 0, st, c, a,
                     b, descs, (uint32_t)nmod);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_polymul(uint64_t *c, const uint64_t *a, const uint64_t *b, const ModDesc *descs,
                          int logn, size_t nres, size_t nmod, NttShape shape, hipStream_t st) {
  if (logn < kMinLogN || logn > kMaxLogN || nres == 0 || nmod == 0 || nres % nmod)
    return hipErrorInvalidValue;
  if (logn > kMaxBlockLog) return hipErrorNotSupported;  // composed by the caller
  const int le = clamp_le(logn, shape.le);
#define YA_PM_CALL(LB, LE) polymul_lb<LB, LE>(c, a, b, descs, nres, nmod, st)
  YA_LB_SWITCH(logn, le, YA_PM_CALL)
#undef YA_PM_CALL
}

}  // namespace ya
