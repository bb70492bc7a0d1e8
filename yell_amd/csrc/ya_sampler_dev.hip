// ya_sampler_dev.hip -- the samplers on the device, writing residue-major
// polynomials straight into HBM (no host sampling, no upload).
//
// Same streams and mappings as the host samplers (ya_sampler.cpp), so a
// device call returns the words of the host call with the same seed:
//   uniform   include/yell/details/poly.hpp:200-228
//   gaussian  include/yell/details/poly.hpp:236-260
//   ZO        include/yell/details/poly.hpp:267-279
//   hwt       include/yell/details/poly.hpp:287-330
// Every kernel is a handful of Philox blocks per thread and one streaming
// write of L x N words: HBM-write bound at the configs' sizes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/yell_amd.h"
#include "params32_table.inc"
#include "params_table.inc"
#include "ya_internal.hpp"
#include "ya_philox.hpp"

namespace ya {
namespace {

__constant__ uint64_t c_primes64[YA_NUM_MODULI] = YA_TABLE_P;
__constant__ uint32_t c_primes32[YA32_NUM_MODULI] = YA32_TABLE_P;

template <class W>
__device__ __forceinline__ uint64_t sampler_prime(uint32_t cm) {
  if constexpr (sizeof(W) == 8) return c_primes64[cm];
  return c_primes32[cm];
}

constexpr unsigned kThreads = 256;
constexpr unsigned kMaxGrid = 16384;

typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));

// Two consecutive words (16-byte aligned for uint64, 8 for uint32).
template <class W>
__device__ __forceinline__ void store2(W *dst, uint64_t a, uint64_t b) {
  if constexpr (sizeof(W) == 8) {
    *(u64x2 *)dst = u64x2{a, b};
  } else {
    dst[0] = (W)a;
    dst[1] = (W)b;
  }
}

// Uniform: one stream block (32 bytes = 32 / sizeof(W) words) per thread
// iteration -- the Philox block (~380 instructions) is the cost at large
// sizes, so each is computed once; a block never straddles two residues
// (degree >= 4 words of 64 bits, >= 8 of 32).  Polynomial blockIdx.y of a launch over several:
// seed + blockIdx.y, at out + blockIdx.y * nblocks * E.
template <class W>
__global__ void __launch_bounds__(kThreads) k_sample_uniform(W *__restrict__ out, uint32_t logn, uint64_t nblocks,
                                                             uint64_t seed) {
  constexpr uint32_t E = 32 / sizeof(W);
  out += (uint64_t)blockIdx.y * nblocks * E;
  seed += blockIdx.y;
  for (uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x; b < nblocks; b += (uint64_t)gridDim.x * kThreads) {
    const Philox4 r = stream_block(seed, b);
    const uint64_t g0 = b * E;
    const uint64_t p = sampler_prime<W>((uint32_t)(g0 >> logn));
    const int top = 63 - __builtin_clzll(p);
    const uint64_t mask = (1ULL << (top + 1)) - 1;  // p < 2^62
    uint64_t v[E];
#pragma unroll
    for (uint32_t e = 0; e < E; ++e) {
      const uint64_t w = (sizeof(W) == 8) ? r.v[e] : (r.v[e >> 1] >> (32 * (e & 1)));
      const uint64_t t = w & mask;
      v[e] = t >= p ? t - p : t;
    }
#pragma unroll
    for (uint32_t e = 0; e < E; e += 2) store2(out + g0 + e, v[e], v[e + 1]);
  }
}

// First k in [0, tail] with u < cdf[k] (cdf non-decreasing, cdf[tail] = 2^64-1):
// the linear scan of the host sampler as a binary search.
__device__ __forceinline__ uint32_t cdt_search(const uint64_t *__restrict__ cdf, uint32_t tail, uint64_t u) {
  uint32_t lo = 0, hi = tail;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (u < cdf[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// a = signed sample times the amplifier in the reference's signed_type
// (details/poly.hpp:248-251: int64_t, or int32_t with 32-bit moduli,
// params_32.hpp), so the product wraps modulo 2^(8 sizeof W); then
// a < 0 ? p + a : a in W (details/poly.hpp:252-257).
template <class W>
__device__ __forceinline__ W gaussian_residue(int64_t s, uint64_t amplifier, uint64_t p) {
  typedef typename std::conditional<sizeof(W) == 8, int64_t, int32_t>::type S;
  const S a = (S)(W)((uint64_t)s * amplifier);
  return a < 0 ? (W)((W)p + (W)a) : (W)a;
}

// Gaussian: one stream block per two coefficients (words 2i, 2i+1 of
// coefficient i), the pair written to every residue.
template <class W>
__global__ void __launch_bounds__(kThreads) k_sample_gaussian(W *__restrict__ out, uint32_t logn, uint32_t nmod,
                                                              const uint64_t *__restrict__ cdf, uint32_t tail,
                                                              uint64_t amplifier, uint64_t seed) {
  const uint64_t npairs = 1ULL << (logn - 1);
  out += ((uint64_t)blockIdx.y * nmod) << logn;
  seed += blockIdx.y;
  for (uint64_t b = (uint64_t)blockIdx.x * kThreads + threadIdx.x; b < npairs; b += (uint64_t)gridDim.x * kThreads) {
    const Philox4 r = stream_block(seed, b);
    const uint32_t k0 = cdt_search(cdf, tail, r.v[0]), k1 = cdt_search(cdf, tail, r.v[2]);
    const int64_t s0 = (k0 != 0 && (r.v[1] >> 63)) ? -(int64_t)k0 : (int64_t)k0;
    const int64_t s1 = (k1 != 0 && (r.v[3] >> 63)) ? -(int64_t)k1 : (int64_t)k1;
    for (uint32_t cm = 0; cm < nmod; ++cm) {
      const uint64_t p = sampler_prime<W>(cm);
      store2(out + ((uint64_t)cm << logn) + 2 * b, gaussian_residue<W>(s0, amplifier, p),
             gaussian_residue<W>(s1, amplifier, p));
    }
  }
}

// ZO: two coefficients (stream bytes 2t, 2t+1) per thread iteration, so a
// wave's stores to a residue cover 1 KiB contiguously (one stream word per
// thread with 64-byte lane strides measured 1.8x slower on MI355X, although
// it computes a quarter of the Philox blocks).
template <class W>
__global__ void __launch_bounds__(kThreads) k_sample_zo(W *__restrict__ out, uint32_t logn, uint32_t nmod, uint32_t rho,
                                                        uint64_t seed) {
  const uint64_t npairs = 1ULL << (logn - 1);
  out += ((uint64_t)blockIdx.y * nmod) << logn;
  seed += blockIdx.y;
  for (uint64_t t = (uint64_t)blockIdx.x * kThreads + threadIdx.x; t < npairs; t += (uint64_t)gridDim.x * kThreads) {
    const uint64_t w = stream_word(seed, t >> 2);
    const uint32_t sh = 16 * (uint32_t)(t & 3);
    const uint32_t r0 = (uint32_t)(w >> sh) & 0xffu, r1 = (uint32_t)(w >> (sh + 8)) & 0xffu;
    for (uint32_t cm = 0; cm < nmod; ++cm) {
      const uint64_t p = sampler_prime<W>(cm);
      const uint64_t v0 = r0 <= rho ? ((r0 & 2) ? p - 1 : 1) : 0;
      const uint64_t v1 = r1 <= rho ? ((r1 & 2) ? p - 1 : 1) : 0;
      store2(out + ((uint64_t)cm << logn) + 2 * t, v0, v1);
    }
  }
}

// ---- hwt: the reservoir of the host sampler, in parallel.  Slot s ends at
// the largest step k >= hwt that drew position s (steps run in increasing k
// on the host, so the last writer is the largest), or at s if none did.
__global__ void __launch_bounds__(kThreads) k_hwt_init(uint32_t *__restrict__ last, uint32_t hwt) {
  for (uint32_t s = blockIdx.x * kThreads + threadIdx.x; s < hwt; s += gridDim.x * kThreads) last[s] = s;
}

__global__ void __launch_bounds__(kThreads) k_fill_u32(uint32_t *__restrict__ p, uint32_t n) {
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) p[i] = 0;
}

__global__ void __launch_bounds__(kThreads) k_hwt_steps(uint32_t *__restrict__ last, uint32_t hwt, uint32_t degree,
                                                        uint64_t seed) {
  for (uint32_t k = hwt + blockIdx.x * kThreads + threadIdx.x; k < degree; k += gridDim.x * kThreads) {
    const uint64_t pos = reservoir_pos(seed, k);
    if (pos < hwt) atomicMax(last + pos, k);
  }
}

__global__ void __launch_bounds__(kThreads) k_hwt_mark(const uint32_t *__restrict__ last, uint32_t *__restrict__ bits,
                                                       uint32_t hwt) {
  for (uint32_t s = blockIdx.x * kThreads + threadIdx.x; s < hwt; s += gridDim.x * kThreads) {
    const uint32_t pos = last[s];
    atomicOr(bits + (pos >> 5), 1u << (pos & 31));
  }
}

// Exclusive prefix of the bitmap's popcounts, one workgroup of 1024 threads.
constexpr unsigned kScanThreads = 1024;
__global__ void __launch_bounds__(kScanThreads) k_hwt_scan(const uint32_t *__restrict__ bits,
                                                           uint32_t *__restrict__ prefix, uint32_t nwords) {
  __shared__ uint32_t part[kScanThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nwords + kScanThreads - 1) / kScanThreads;
  const uint32_t w0 = t * per, w1 = min(nwords, w0 + per);
  uint32_t sum = 0;
  for (uint32_t w = w0; w < w1; ++w) sum += __popc(bits[w]);
  part[t] = sum;
  __syncthreads();
  for (uint32_t d = 1; d < kScanThreads; d <<= 1) {  // inclusive Hillis-Steele scan
    const uint32_t add = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (uint32_t w = w0; w < w1; ++w) {
    prefix[w] = run;
    run += __popc(bits[w]);
  }
}

// The hit of rank j (j-th smallest position) gets +1 if stream word j has
// bit 1 set, else p-1.
template <class W>
__global__ void __launch_bounds__(kThreads) k_hwt_write(W *__restrict__ out, const uint32_t *__restrict__ bits,
                                                        const uint32_t *__restrict__ prefix, uint32_t logn,
                                                        uint32_t nmod, uint64_t seed) {
  const uint32_t n = 1u << logn;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const uint32_t word = bits[i >> 5];
    const bool hit = (word >> (i & 31)) & 1u;
    bool plus = false;
    if (hit) {
      const uint32_t rank = prefix[i >> 5] + __popc(word & ((1u << (i & 31)) - 1u));
      plus = (stream_word(seed, rank) & 2) != 0;
    }
    for (uint32_t cm = 0; cm < nmod; ++cm) {
      const uint64_t p = sampler_prime<W>(cm);
      out[((uint64_t)cm << logn) + i] = (W)(hit ? (plus ? 1 : p - 1) : 0);
    }
  }
}

unsigned grid_for(uint64_t items) {
  const uint64_t g = (items + kThreads - 1) / kThreads;
  return (unsigned)(g == 0 ? 1 : (g > kMaxGrid ? kMaxGrid : g));
}

// Device copies of the Gaussian tables, one per (device, sigma), kept for the
// life of the process.
constexpr size_t kMaxCachedSigmas = 64;
std::mutex g_cdt_mu;
std::map<std::pair<int, uint64_t>, std::pair<uint64_t *, uint32_t>> g_cdt;

int device_cdt(double sigma, const uint64_t **table, uint32_t *tail) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  uint64_t bits;
  std::memcpy(&bits, &sigma, sizeof bits);
  std::lock_guard<std::mutex> lk(g_cdt_mu);
  auto it = g_cdt.find({dev, bits});
  if (it == g_cdt.end()) {
    if (g_cdt.size() >= kMaxCachedSigmas)
      return fail(YA_ERR_INVALID, "more than %zu distinct Gaussian parameters", kMaxCachedSigmas);
    std::vector<uint64_t> cdf;
    int rc = gaussian_cdt(sigma, cdf);
    if (rc) return rc;
    uint64_t *d = nullptr;
    e = hipMalloc(&d, cdf.size() * sizeof(uint64_t));
    if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipMalloc(gaussian table): %s", hipGetErrorString(e));
    e = hipMemcpy(d, cdf.data(), cdf.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return hip_fail(e, "gaussian table upload");
    }
    it = g_cdt.emplace(std::make_pair(dev, bits), std::make_pair(d, (uint32_t)(cdf.size() - 1))).first;
  }
  *table = it->second.first;
  *tail = it->second.second;
  return YA_OK;
}

// Shape and pointer checks of either build.
template <class W>
int dev_shape(const W *out, size_t degree, size_t nmoduli, int *logn) {
  if (!out) return fail(YA_ERR_INVALID, "null output");
  if (((uintptr_t)out) % 16) return fail(YA_ERR_INVALID, "output must be 16-byte aligned");
  if constexpr (sizeof(W) == 8) {
    return check_shape(degree, 0, nmoduli, logn);
  } else {
    const int l = log2_degree(degree);
    if (l < 3 || l > 15) return fail(YA_ERR_INVALID, "degree %zu: a power of two in [8, 2^15] is required", degree);
    if (nmoduli == 0 || nmoduli > YA32_NUM_MODULI)
      return fail(YA_ERR_INVALID, "nmoduli %zu out of [1, %d]", nmoduli, YA32_NUM_MODULI);
    *logn = l;
    return YA_OK;
  }
}

int launched(const char *what) {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? YA_OK : hip_fail(e, what);
}

// Polynomials are launched in chunks of at most 65535 (grid.y).
constexpr size_t kMaxPolysPerLaunch = 65535;

int check_count(size_t count, size_t nmoduli) {
  if (count == 0 || count * nmoduli > (1u << 30)) return fail(YA_ERR_INVALID, "bad count %zu", count);
  return YA_OK;
}

template <class W>
int dev_uniform(W *out, size_t degree, size_t nmoduli, size_t count, uint64_t seed, void *stream) {
  int logn;
  int rc = dev_shape(out, degree, nmoduli, &logn);
  if (!rc) rc = check_count(count, nmoduli);
  if (rc) return rc;
  const uint64_t nblocks = (uint64_t)degree * nmoduli * sizeof(W) / 32;
  for (size_t k0 = 0; k0 < count; k0 += kMaxPolysPerLaunch) {
    const size_t kc = std::min(kMaxPolysPerLaunch, count - k0);
    hipLaunchKernelGGL(k_sample_uniform<W>, dim3(grid_for(nblocks), (unsigned)kc), dim3(kThreads), 0,
                       (hipStream_t)stream, out + k0 * degree * nmoduli, (uint32_t)logn, nblocks, seed + k0);
  }
  return launched("uniform sampler launch");
}

template <class W>
int dev_gaussian(W *out, size_t degree, size_t nmoduli, size_t count, double sigma, uint64_t amplifier,
                 uint64_t seed, void *stream) {
  int logn;
  int rc = dev_shape(out, degree, nmoduli, &logn);
  if (!rc) rc = check_count(count, nmoduli);
  if (rc) return rc;
  const uint64_t *cdf;
  uint32_t tail;
  rc = device_cdt(sigma, &cdf, &tail);
  if (rc) return rc;
  for (size_t k0 = 0; k0 < count; k0 += kMaxPolysPerLaunch) {
    const size_t kc = std::min(kMaxPolysPerLaunch, count - k0);
    hipLaunchKernelGGL(k_sample_gaussian<W>, dim3(grid_for(degree / 2), (unsigned)kc), dim3(kThreads), 0,
                       (hipStream_t)stream, out + k0 * degree * nmoduli, (uint32_t)logn, (uint32_t)nmoduli, cdf,
                       tail, amplifier, seed + k0);
  }
  return launched("gaussian sampler launch");
}

template <class W>
int dev_zo(W *out, size_t degree, size_t nmoduli, size_t count, uint32_t rho, uint64_t seed, void *stream) {
  if (rho > 255) return fail(YA_ERR_INVALID, "rho %u > 255", rho);
  int logn;
  int rc = dev_shape(out, degree, nmoduli, &logn);
  if (!rc) rc = check_count(count, nmoduli);
  if (rc) return rc;
  for (size_t k0 = 0; k0 < count; k0 += kMaxPolysPerLaunch) {
    const size_t kc = std::min(kMaxPolysPerLaunch, count - k0);
    hipLaunchKernelGGL(k_sample_zo<W>, dim3(grid_for(degree / 2), (unsigned)kc), dim3(kThreads), 0,
                       (hipStream_t)stream, out + k0 * degree * nmoduli, (uint32_t)logn, (uint32_t)nmoduli, rho,
                       seed + k0);
  }
  return launched("zo sampler launch");
}

template <class W>
int dev_hwt(W *out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed, void *stream) {
  int logn;
  int rc = dev_shape(out, degree, nmoduli, &logn);
  if (rc) return rc;
  if (hwt == 0 || hwt > degree) return fail(YA_ERR_INVALID, "hwt must be in [1, degree]");
  const hipStream_t s = (hipStream_t)stream;
  const uint32_t nwords = (uint32_t)((degree + 31) / 32);
  const size_t words = (size_t)hwt + 2 * (size_t)nwords;
  uint32_t *scratch = nullptr;
  hipError_t e = hipMallocAsync((void **)&scratch, words * sizeof(uint32_t), s);
  if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipMallocAsync(hwt scratch): %s", hipGetErrorString(e));
  uint32_t *last = scratch, *bits = scratch + hwt, *prefix = bits + nwords;
  const uint32_t n = (uint32_t)degree;
  hipLaunchKernelGGL(k_hwt_init, dim3(grid_for(hwt)), dim3(kThreads), 0, s, last, hwt);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nwords)), dim3(kThreads), 0, s, bits, nwords);
  if (hwt < n) hipLaunchKernelGGL(k_hwt_steps, dim3(grid_for(n - hwt)), dim3(kThreads), 0, s, last, hwt, n, seed);
  hipLaunchKernelGGL(k_hwt_mark, dim3(grid_for(hwt)), dim3(kThreads), 0, s, last, bits, hwt);
  hipLaunchKernelGGL(k_hwt_scan, dim3(1), dim3(kScanThreads), 0, s, bits, prefix, nwords);
  hipLaunchKernelGGL(k_hwt_write<W>, dim3(grid_for(n)), dim3(kThreads), 0, s, out, bits, prefix, (uint32_t)logn,
                     (uint32_t)nmoduli, seed);
  rc = launched("hwt sampler launch");
  e = hipFreeAsync(scratch, s);
  if (!rc && e != hipSuccess) rc = hip_fail(e, "hipFreeAsync(hwt scratch)");
  return rc;
}

}  // namespace
}  // namespace ya

using namespace ya;

extern "C" {

int ya_sample_uniform_device(uint64_t *d_out, size_t degree, size_t nmoduli, uint64_t seed, void *stream) {
  return dev_uniform(d_out, degree, nmoduli, 1, seed, stream);
}
int ya_sample_gaussian_device(uint64_t *d_out, size_t degree, size_t nmoduli, double sigma, uint64_t amplifier,
                              uint64_t seed, void *stream) {
  return dev_gaussian(d_out, degree, nmoduli, 1, sigma, amplifier, seed, stream);
}
int ya_sample_zo_device(uint64_t *d_out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed, void *stream) {
  return dev_zo(d_out, degree, nmoduli, 1, rho, seed, stream);
}
int ya_sample_hwt_device(uint64_t *d_out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed, void *stream) {
  return dev_hwt(d_out, degree, nmoduli, hwt, seed, stream);
}
int ya_sample_uniform_device_many(uint64_t *d_out, size_t degree, size_t nmoduli, size_t count, uint64_t seed,
                                  void *stream) {
  return dev_uniform(d_out, degree, nmoduli, count, seed, stream);
}
int ya_sample_gaussian_device_many(uint64_t *d_out, size_t degree, size_t nmoduli, size_t count, double sigma,
                                   uint64_t amplifier, uint64_t seed, void *stream) {
  return dev_gaussian(d_out, degree, nmoduli, count, sigma, amplifier, seed, stream);
}
int ya_sample_zo_device_many(uint64_t *d_out, size_t degree, size_t nmoduli, size_t count, uint32_t rho,
                             uint64_t seed, void *stream) {
  return dev_zo(d_out, degree, nmoduli, count, rho, seed, stream);
}

int ya32_sample_uniform_device(uint32_t *d_out, size_t degree, size_t nmoduli, uint64_t seed, void *stream) {
  return dev_uniform(d_out, degree, nmoduli, 1, seed, stream);
}
int ya32_sample_gaussian_device(uint32_t *d_out, size_t degree, size_t nmoduli, double sigma, uint64_t amplifier,
                                uint64_t seed, void *stream) {
  return dev_gaussian(d_out, degree, nmoduli, 1, sigma, amplifier, seed, stream);
}
int ya32_sample_zo_device(uint32_t *d_out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed, void *stream) {
  return dev_zo(d_out, degree, nmoduli, 1, rho, seed, stream);
}
int ya32_sample_hwt_device(uint32_t *d_out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed, void *stream) {
  return dev_hwt(d_out, degree, nmoduli, hwt, seed, stream);
}

}  // extern "C"
