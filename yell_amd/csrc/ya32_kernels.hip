// ya32_kernels.hip -- kernels of the 32-bit moduli build
// (YELL_USE_32BITS_MODULI: include/yell/params_32.hpp; uint32 words, primes
// 2^30 - k*2^16 + 1, degrees up to 2^15).
//
// The same pass structure as the 64-bit kernels (ya_blocks.hpp, generic in
// the word type): a residue of 2^LB <= 2^15 words is one workgroup, 2^LE
// words per thread, LE layers per register pass, LDS exchanges between
// passes (2^15 words * 4 B + padding = 136 KiB).  Each butterfly is one
// v_mul_hi_u32 and two v_mul_lo_u32 (lib/ntt.cpp:47-83 with bit_width 32)
// instead of the 64-bit build's ~10 multiply-class instructions.
#include "ya32.hpp"
#include "ya_blocks.hpp"

namespace ya {

template <int LB, int LE, bool LAZY>
__global__ void __launch_bounds__(1 << (LB - LE))
k32_fwd(uint32_t *__restrict__ data, const ModDesc32 *__restrict__ descs, uint32_t nmod) {
  __shared__ uint32_t lds[LdsWords<LB>::value];
  const uint32_t res = blockIdx.x;
  const ModDesc32 d = descs[res % nmod];
  uint32_t *base = data + ((size_t)res << LB);
  const uint32_t t = threadIdx.x;
  uint32_t x[1 << LE];
  fwd_block<LB, LE>(x, base, lds, d.fwd, LB, 0u, d.p, t);
  if constexpr (!LAZY) correct_all<LE>(x, d.p);
  store_coalesced<LB, LE>(base, x, lds, t);
}

template <int LB, int LE>
__global__ void __launch_bounds__(1 << (LB - LE))
k32_inv(uint32_t *__restrict__ data, const ModDesc32 *__restrict__ descs, uint32_t nmod) {
  __shared__ uint32_t lds[LdsWords<LB>::value];
  const uint32_t res = blockIdx.x;
  const ModDesc32 d = descs[res % nmod];
  uint32_t *base = data + ((size_t)res << LB);
  const uint32_t t = threadIdx.x;
  uint32_t x[1 << LE];
  load_coalesced<LB, LE>(base, x, lds, t);
  inv_block<LB, LE, TOP_MERGED>(x, base, lds, d.inv, LB, 0u, d, t);
}

// Fused c = backward(forward(a) (*) forward(b)) of one residue per workgroup;
// residue order prime-major over the XCDs, as the 64-bit k_polymul.
template <int LB, int LE, bool DUAL>
__device__ __forceinline__ void polymul32(uint32_t *__restrict__ c, const uint32_t *__restrict__ a,
                                          const uint32_t *__restrict__ b, const ModDesc32 *__restrict__ descs,
                                          uint32_t nmod, uint32_t *lds) {
  const uint32_t g = gridDim.x;
  const uint32_t v = xcd_order(blockIdx.x, g);
  const uint32_t count = g / nmod;
  const uint32_t cm = v / count, poly = v % count;
  const uint32_t res = poly * nmod + cm;
  const ModDesc32 d = descs[cm];
  const size_t off = (size_t)res << LB;
  uint32_t *cr = uniform_ptr(c + off);
  const uint32_t *ar = uniform_ptr(a + off), *br = uniform_ptr(b + off);
  const uint32_t t = threadIdx.x;
  uint32_t x[1 << LE], y[1 << LE];
  constexpr int W0 = FirstWindow<LB, LE>::value;
  gmem_load<W0, LE>(ar, x, t);
  gmem_load<W0, LE>(br, y, t);
  if constexpr (DUAL) {
    // both forward transforms in one pass sequence (shared twiddles)
    fwd_block_dual<LB, LE>(x, y, lds, (const TwC32 *)d.fwd, LB, 0u, d.p, t);
  } else {
    fwd_block<LB, LE, true>(x, nullptr, lds, d.fwd, LB, 0u, d.p, t);
    gmem_store<0, LE>(cr, x, t);  // park forward(a) at this thread's own words
#pragma unroll
    for (int e = 0; e < (1 << LE); ++e) x[e] = y[e];
    const uint32_t t2 = phase_v(t);
    fwd_block<LB, LE, true>(x, nullptr, lds, d.fwd + phase_zero_s(), LB, 0u, d.p, t2);
    gmem_load<0, LE>(cr, y, t2);
  }
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(y[e], x[e], d.p, d.pn);
  const uint32_t t3 = phase_v(t);
  inv_block<LB, LE, TOP_MERGED>(x, cr, lds, (const TwC32 *)d.inv, LB, 0u, d, t3);
}

template <int LB, int LE, bool DUAL>
__global__ void __launch_bounds__(1 << (LB - LE))
k32_polymul(uint32_t *__restrict__ c, const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
            const ModDesc32 *__restrict__ descs, uint32_t nmod) {
  __shared__ uint32_t lds[LdsWords<LB>::value];
  polymul32<LB, LE, DUAL>(c, a, b, descs, nmod, lds);
}

// The same at 2^14 built for two workgroups per CU: a residue takes 68 KiB of
// LDS, so two 1024-thread workgroups fit when each wave keeps to 64 VGPRs
// (8 waves per SIMD), at the price of a few bytes of spills.  Measured on 256
// poly-mults of 16 residues: 0.45 ms at one workgroup per CU, 0.40 ms at two.
template <int LB, int LE, bool DUAL>
__global__ void __launch_bounds__(1 << (LB - LE)) __attribute__((amdgpu_waves_per_eu(8)))
k32_polymul_2wg(uint32_t *__restrict__ c, const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                const ModDesc32 *__restrict__ descs, uint32_t nmod) {
  __shared__ uint32_t lds[LdsWords<LB>::value];
  polymul32<LB, LE, DUAL>(c, a, b, descs, nmod, lds);
}

// Element-wise operations over residue-major words (include/yell/ops.hpp:7-41,
// details/poly.hpp:119-158); four words per thread.
template <int OP>
__global__ void __launch_bounds__(256)
k32_pointwise(uint32_t *__restrict__ out, const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
              const ModDesc32 *__restrict__ descs, uint32_t logn, uint32_t nmod, size_t words) {
  const size_t i0 = ((size_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i0 >= words) return;
  const uint32_t res = (uint32_t)(i0 >> logn);
  const ModDesc32 d = descs[res % nmod];
  const size_t n = words - i0 < 4 ? words - i0 : 4;  // 4 words never straddle a residue (N >= 8)
  for (size_t k = 0; k < n; ++k) {
    const size_t i = i0 + k;
    uint32_t r;
    if constexpr (OP == PW_MUL) r = mulmod_barrett(a[i], b[i], d.p, d.pn);
    else if constexpr (OP == PW_ADD) r = addmod(a[i], b[i], d.p);
    else if constexpr (OP == PW_SUB) r = submod(a[i], b[i], d.p);
    else if constexpr (OP == PW_NEGATE) r = d.p - a[i];
    else r = muladd_barrett(out[i], a[i], b[i], d.p, d.pn);  // PW_MULADD: out += a*b
    out[i] = r;
  }
}

namespace {

// (lb, le) shapes: le 3 at 2^3, 4 for 2^4 .. 2^14, 5 at 2^15 (1024 threads).
#define YA32_SWITCH(lb, CALL)   \
  switch (lb) {                 \
    case 3: return CALL(3, 3);  \
    case 4: return CALL(4, 4);  \
    case 5: return CALL(5, 4);  \
    case 6: return CALL(6, 4);  \
    case 7: return CALL(7, 4);  \
    case 8: return CALL(8, 4);  \
    case 9: return CALL(9, 4);  \
    case 10: return CALL(10, 4); \
    case 11: return CALL(11, 4); \
    case 12: return CALL(12, 4); \
    case 13: return CALL(13, 4); \
    case 14: return CALL(14, 4); \
    case 15: return CALL(15, 5); \
    default: return hipErrorInvalidValue; \
  }

template <int LB, int LE>
hipError_t fwd32(uint32_t *data, const ModDesc32 *descs, size_t nres, size_t nmod, bool lazy, hipStream_t st) {
  if (lazy)
    hipLaunchKernelGGL((k32_fwd<LB, LE, true>), dim3((unsigned)nres), dim3(1u << (LB - LE)), 0, st, data, descs,
                       (uint32_t)nmod);
  else
    hipLaunchKernelGGL((k32_fwd<LB, LE, false>), dim3((unsigned)nres), dim3(1u << (LB - LE)), 0, st, data, descs,
                       (uint32_t)nmod);
  return hipGetLastError();
}

template <int LB, int LE>
hipError_t inv32(uint32_t *data, const ModDesc32 *descs, size_t nres, size_t nmod, hipStream_t st) {
  hipLaunchKernelGGL((k32_inv<LB, LE>), dim3((unsigned)nres), dim3(1u << (LB - LE)), 0, st, data, descs,
                     (uint32_t)nmod);
  return hipGetLastError();
}

// The dual-forward form from blocks of 2^10 on (as the 64-bit build); below,
// forward(a) parks in c.
template <int LB, int LE>
hipError_t pm32(uint32_t *c, const uint32_t *a, const uint32_t *b, const ModDesc32 *descs, size_t nres,
                size_t nmod, hipStream_t st) {
  if constexpr (LB == 14 && LE == 4)
    hipLaunchKernelGGL((k32_polymul_2wg<LB, LE, true>), dim3((unsigned)nres), dim3(1u << (LB - LE)), 0, st, c, a,
                       b, descs, (uint32_t)nmod);
  else
    hipLaunchKernelGGL((k32_polymul<LB, LE, (LB >= 10)>), dim3((unsigned)nres), dim3(1u << (LB - LE)), 0, st, c, a,
                       b, descs, (uint32_t)nmod);
  return hipGetLastError();
}

}  // namespace

hipError_t launch32_forward(uint32_t *data, const ModDesc32 *descs, int logn, size_t nres, size_t nmod,
                            bool lazy, hipStream_t st) {
  if (nres == 0 || nmod == 0 || nres > 0x7fffffff) return hipErrorInvalidValue;
#define YA32_FWD(LB, LE) fwd32<LB, LE>(data, descs, nres, nmod, lazy, st)
  YA32_SWITCH(logn, YA32_FWD)
#undef YA32_FWD
}

hipError_t launch32_backward(uint32_t *data, const ModDesc32 *descs, int logn, size_t nres, size_t nmod,
                             hipStream_t st) {
  if (nres == 0 || nmod == 0 || nres > 0x7fffffff) return hipErrorInvalidValue;
#define YA32_INV(LB, LE) inv32<LB, LE>(data, descs, nres, nmod, st)
  YA32_SWITCH(logn, YA32_INV)
#undef YA32_INV
}

hipError_t launch32_polymul(uint32_t *c, const uint32_t *a, const uint32_t *b, const ModDesc32 *descs,
                            int logn, size_t nres, size_t nmod, hipStream_t st) {
  if (nres == 0 || nmod == 0 || nres % nmod || nres > 0x7fffffff) return hipErrorInvalidValue;
#define YA32_PM(LB, LE) pm32<LB, LE>(c, a, b, descs, nres, nmod, st)
  YA32_SWITCH(logn, YA32_PM)
#undef YA32_PM
}

hipError_t launch32_pointwise(PwOp op, uint32_t *out, const uint32_t *a, const uint32_t *b,
                              const ModDesc32 *descs, int logn, size_t nmod, size_t words, hipStream_t st) {
  if (words == 0) return hipSuccess;
  const size_t blocks = (words + 1023) / 1024;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  const dim3 g((unsigned)blocks), blk(256);
  switch (op) {
    case PW_MUL: hipLaunchKernelGGL(k32_pointwise<PW_MUL>, g, blk, 0, st, out, a, b, descs, (uint32_t)logn, (uint32_t)nmod, words); break;
    case PW_ADD: hipLaunchKernelGGL(k32_pointwise<PW_ADD>, g, blk, 0, st, out, a, b, descs, (uint32_t)logn, (uint32_t)nmod, words); break;
    case PW_SUB: hipLaunchKernelGGL(k32_pointwise<PW_SUB>, g, blk, 0, st, out, a, b, descs, (uint32_t)logn, (uint32_t)nmod, words); break;
    case PW_NEGATE: hipLaunchKernelGGL(k32_pointwise<PW_NEGATE>, g, blk, 0, st, out, a, b, descs, (uint32_t)logn, (uint32_t)nmod, words); break;
    case PW_MULADD: hipLaunchKernelGGL(k32_pointwise<PW_MULADD>, g, blk, 0, st, out, a, b, descs, (uint32_t)logn, (uint32_t)nmod, words); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ya
