// ya_probe.hip -- VALU throughput probes used as the compute ceiling of the
// roofline: the rate at which the device retires 32-bit integer multiplies
// and this engine's exact CT butterflies (64-bit words, and the 32-bit moduli
// build's) with all operands in registers (no memory traffic inside the timed
// loop).
#include <hip/hip_runtime.h>

#include "../../include/yell_amd.h"
#include "ya_internal.hpp"

namespace ya {

constexpr int kProbeThreads = 256;
constexpr int kChains = 8;

template <int KIND>
__global__ void __launch_bounds__(kProbeThreads)
k_probe(uint64_t *__restrict__ out, uint32_t iters, uint64_t seed, uint64_t p) {
  const uint32_t gid = blockIdx.x * kProbeThreads + threadIdx.x;
  uint64_t v[kChains * 2];
#pragma unroll
  for (int k = 0; k < kChains * 2; ++k) v[k] = (seed ^ (gid * 0x9e3779b97f4a7c15ULL)) + k * 0x1234567ULL;
  const uint32_t m = (uint32_t)(seed >> 7) | 1u;
  if constexpr (KIND == 0) {  // v_mad_u64_u32
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
      for (int k = 0; k < kChains; ++k) v[k] = (uint64_t)(uint32_t)v[k] * m + v[k];
    }
  } else if constexpr (KIND == 1) {  // v_mul_lo_u32
    uint32_t w[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) w[k] = (uint32_t)v[k];
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
      for (int k = 0; k < kChains; ++k) w[k] = w[k] * m;
    }
#pragma unroll
    for (int k = 0; k < kChains; ++k) v[k] = w[k];
  } else if constexpr (KIND == 3) {  // 32-bit moduli build: CT butterflies over 16 words
    const uint32_t q = (uint32_t)p;
    const PrimeK32 pk(q);
    const uint32_t w0 = (uint32_t)(seed % q);
    const Tw32 t{w0, (uint32_t)((((uint64_t)w0) << 32) / q)};
    uint32_t x[kChains * 2];
#pragma unroll
    for (int k = 0; k < kChains * 2; ++k) x[k] = (uint32_t)(v[k] % ((uint64_t)q << 2));
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
      for (int k = 0; k < kChains; ++k) ct_butterfly(x[k], x[k + kChains], t, pk);
    }
#pragma unroll
    for (int k = 0; k < kChains * 2; ++k) v[k] = x[k];
  } else {  // CT butterflies over 16 words, 8 per iteration
    const PrimeK pk(p);
    const Tw t{(seed % p), ((seed % p) * 3) | 1};
#pragma unroll
    for (int k = 0; k < kChains * 2; ++k) v[k] %= (p << 2);
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
      for (int k = 0; k < kChains; ++k) ct_butterfly(v[k], v[k + kChains], t, pk);
    }
  }
  uint64_t acc = 0;
#pragma unroll
  for (int k = 0; k < kChains * 2; ++k) acc ^= v[k];
  out[gid] = acc;
}

}  // namespace ya

using namespace ya;

extern "C" int ya_probe_valu(int kind, double *ops_per_second) {
  if (!ops_per_second || kind < 0 || kind > 3) return fail(YA_ERR_INVALID, "bad probe kind");
  DeviceState *st;
  int rc = current_device(&st);
  if (rc) return rc;
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return hip_fail(e, "device attributes");
  const unsigned blocks = (unsigned)cus * 16u;
  const uint32_t iters = kind >= 2 ? 2048u : 8192u;
  uint64_t *out = nullptr;
  e = hipMalloc(&out, (size_t)blocks * kProbeThreads * 8);
  if (e != hipSuccess) return fail(YA_ERR_ALLOC, "probe buffer");
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const uint64_t p = kind == 3 ? (uint64_t)ya32_prime(0) : prime(0);
  float best = 1e30f;
  for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
    (void)hipEventRecord(a, nullptr);
    switch (kind) {
      case 0: hipLaunchKernelGGL(k_probe<0>, dim3(blocks), dim3(kProbeThreads), 0, nullptr, out, iters, 0x5eedULL + rep, p); break;
      case 1: hipLaunchKernelGGL(k_probe<1>, dim3(blocks), dim3(kProbeThreads), 0, nullptr, out, iters, 0x5eedULL + rep, p); break;
      case 2: hipLaunchKernelGGL(k_probe<2>, dim3(blocks), dim3(kProbeThreads), 0, nullptr, out, iters, 0x5eedULL + rep, p); break;
      default: hipLaunchKernelGGL(k_probe<3>, dim3(blocks), dim3(kProbeThreads), 0, nullptr, out, iters, 0x5eedULL + rep, p); break;
    }
    e = hipGetLastError();
    (void)hipEventRecord(b, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (ms < best) best = ms;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(out);
  if (e != hipSuccess) return hip_fail(e, "probe");
  const double ops = (double)blocks * kProbeThreads * iters * kChains;
  *ops_per_second = ops / (best * 1e-3);
  return YA_OK;
}
