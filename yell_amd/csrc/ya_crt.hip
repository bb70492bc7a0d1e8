// ya_crt.hip -- CRT lift of RNS residues to multi-limb integers on the GPU.
//
// Follows gmp::poly2mpz (include/yell/gmp.hpp:41-70) with the constants of
// gmp_precomputed (lib/gmp.cpp:6-51), per coefficient i of L residues r_j:
//   S   = sum_j lift_j * r_j                  (lift_j = (Q/p_j) * ((Q/p_j)^-1 mod p_j))
//   tmp = floor(S * M / 2^k)                  (M = floor(2^k / Q),
//                                              k = bits(Q) + 64 + floor(log2 L) + 1)
//   S  -= tmp * Q;  S -= Q if S >= Q
// all in exact integer arithmetic, as GMP computes it.  With k that large the
// quotient estimate is off by at most one for any 64-bit residues, so the
// result is S mod Q, in [0, Q): ceil(bits(Q) / 64) limbs per coefficient.
//
// Layout: residues [L][degree] (residue-major, as every other entry point);
// output [degree][limbs] little-endian 64-bit limbs per coefficient, the word
// order mpz_import(rop, limbs, -1, 8, 0, 0, ptr) reads.
//
// One thread per coefficient.  Up to 32 primes the running sum lives in
// registers (32-bit limbs, v_mad_u64_u32 chains, lift limbs as wave-uniform
// scalar loads); beyond that a limb-major scratch buffer holds it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/yell_amd.h"
#include "ya_internal.hpp"

namespace ya {

namespace {

constexpr int kCrtThreads = 256;
constexpr int kCrtMaxRegPrimes = 32;     // register path up to this many primes
constexpr size_t kCrtChunk = 8192;       // coefficients per launch on the scratch path

// ---------------------------------------------------------------- host side

using u128 = unsigned __int128;
using Big = std::vector<uint64_t>;  // little-endian 64-bit limbs

size_t big_bits(const Big &v) {
  for (size_t k = v.size(); k-- > 0;)
    if (v[k]) return 64 * k + 64 - (size_t)__builtin_clzll(v[k]);
  return 0;
}

void big_mul_small(Big &v, uint64_t m) {
  u128 c = 0;
  for (auto &x : v) {
    c += (u128)x * m;
    x = (uint64_t)c;
    c >>= 64;
  }
  if (c) v.push_back((uint64_t)c);
}

// v /= d (exact or not); returns the remainder.
uint64_t big_div_small(Big &v, uint64_t d) {
  u128 r = 0;
  for (size_t k = v.size(); k-- > 0;) {
    r = (r << 64) | v[k];
    v[k] = (uint64_t)(r / d);
    r %= d;
  }
  while (v.size() > 1 && v.back() == 0) v.pop_back();
  return (uint64_t)r;
}

uint64_t big_mod_small(const Big &v, uint64_t d) {
  u128 r = 0;
  for (size_t k = v.size(); k-- > 0;) r = ((r << 64) | v[k]) % d;
  return (uint64_t)r;
}

uint64_t powmod(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p;
  a %= p;
  while (e) {
    if (e & 1) r = (uint64_t)((u128)r * a % p);
    a = (uint64_t)((u128)a * a % p);
    e >>= 1;
  }
  return r;
}

// a >= b over n limbs (both sized n)
bool big_ge(const Big &a, const Big &b) {
  for (size_t k = a.size(); k-- > 0;)
    if (a[k] != b[k]) return a[k] > b[k];
  return true;
}

void big_sub_inplace(Big &a, const Big &b) {
  uint64_t br = 0;
  for (size_t k = 0; k < a.size(); ++k) {
    const uint64_t bk = k < b.size() ? b[k] : 0;
    const u128 t = (u128)a[k] - bk - br;
    a[k] = (uint64_t)t;
    br = (uint64_t)(t >> 64) ? 1 : 0;
  }
}

void big_shl1(Big &a) {
  uint64_t c = 0;
  for (auto &x : a) {
    const uint64_t nc = x >> 63;
    x = (x << 1) | c;
    c = nc;
  }
}

// Constants of lib/gmp.cpp:6-51 for the primes P[0..L).
struct CrtHost {
  size_t nmoduli = 0;
  size_t bits = 0;      // bits_in_moduli_product
  size_t shift = 0;     // shift_modulus_shoup
  size_t limbs = 0;     // ceil(bits / 64): limbs of a lifted coefficient
  Big q;                // moduli product, `limbs` limbs
  u128 m = 0;           // modulus_shoup = floor(2^shift / Q) (< 2^75)
  std::vector<Big> lift;  // lifting_integers, each `limbs` limbs
};

size_t floor_log2(size_t v) {
  size_t r = 0;
  while (v >>= 1) ++r;
  return r;
}

std::unique_ptr<CrtHost> build_crt_host(size_t L) {
  auto h = std::make_unique<CrtHost>();
  h->nmoduli = L;
  Big q{1};
  for (size_t j = 0; j < L; ++j) big_mul_small(q, prime(j));
  h->bits = big_bits(q);
  h->limbs = (h->bits + 63) / 64;
  q.resize(h->limbs, 0);
  h->q = q;
  // params::kModulusRepresentationBitsize = 64; (T)std::log2(L) = floor(log2 L)
  h->shift = h->bits + 64 + floor_log2(L) + 1;
  // floor(2^shift / Q) by restoring division over the quotient bits: after the
  // top `bits` dividend bits the partial remainder is 2^(bits-1) < Q.
  Big r(h->limbs + 1, 0), qq = q;
  qq.push_back(0);
  r[(h->bits - 1) / 64] = 1ull << ((h->bits - 1) % 64);
  u128 m = 0;
  for (size_t pos = h->shift - h->bits + 1; pos-- > 0;) {
    big_shl1(r);
    m <<= 1;
    if (big_ge(r, qq)) {
      big_sub_inplace(r, qq);
      m |= 1;
    }
  }
  h->m = m;
  h->lift.resize(L);
  for (size_t j = 0; j < L; ++j) {
    const uint64_t p = prime(j);
    Big quo = q;
    (void)big_div_small(quo, p);  // exact
    const uint64_t inv = powmod(big_mod_small(quo, p), p - 2, p);
    big_mul_small(quo, inv);
    quo.resize(h->limbs, 0);  // < Q
    h->lift[j] = quo;
  }
  return h;
}

std::mutex g_crt_mu;
std::map<size_t, std::unique_ptr<CrtHost>> g_crt_host;

const CrtHost *crt_host(size_t L) {
  std::lock_guard<std::mutex> g(g_crt_mu);
  auto &slot = g_crt_host[L];
  if (!slot) slot = build_crt_host(L);
  return slot.get();
}

// Device copy: 32-bit limbs, Q and each lift_j padded to wq limbs.
struct CrtDev {
  int wq = 0;                   // 32-bit limbs per lift / Q on the device
  uint32_t *lift = nullptr;     // [L][wq]
  uint32_t *q = nullptr;        // [wq]
  uint32_t m[3] = {0, 0, 0};    // modulus_shoup, 32-bit limbs
};

std::map<std::pair<int, size_t>, std::unique_ptr<CrtDev>> g_crt_dev;

int reg_bucket(size_t wq_needed) {
  static const int kBuckets[] = {4, 8, 16, 24, 32, 48, 64};
  for (int b : kBuckets)
    if ((size_t)b >= wq_needed) return b;
  return 0;
}

int crt_device(size_t L, const CrtHost *h, const CrtDev **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> g(g_crt_mu);
  auto &slot = g_crt_dev[{dev, L}];
  if (!slot) {
    auto d = std::make_unique<CrtDev>();
    const size_t need = 2 * h->limbs;
    const int b = L <= (size_t)kCrtMaxRegPrimes ? reg_bucket(need) : 0;
    d->wq = b ? b : (int)need;
    std::vector<uint32_t> lift((size_t)L * d->wq, 0), q(d->wq, 0);
    for (size_t j = 0; j < L; ++j)
      for (size_t k = 0; k < need; ++k) lift[j * d->wq + k] = (uint32_t)(h->lift[j][k / 2] >> (32 * (k & 1)));
    for (size_t k = 0; k < need; ++k) q[k] = (uint32_t)(h->q[k / 2] >> (32 * (k & 1)));
    d->m[0] = (uint32_t)h->m;
    d->m[1] = (uint32_t)(h->m >> 32);
    d->m[2] = (uint32_t)(h->m >> 64);
    e = hipMalloc(&d->lift, lift.size() * 4);
    if (e == hipSuccess) e = hipMalloc(&d->q, q.size() * 4);
    if (e != hipSuccess) {
      (void)hipFree(d->lift);
      return fail(YA_ERR_ALLOC, "crt tables: %s", hipGetErrorString(e));
    }
    e = hipMemcpy(d->lift, lift.data(), lift.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d->q, q.data(), q.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d->lift);
      (void)hipFree(d->q);
      return hip_fail(e, "crt tables upload");
    }
    slot = std::move(d);
  }
  *out = slot.get();
  return YA_OK;
}

}  // namespace

// -------------------------------------------------------------- device side

// acc[off..NS) += a[0..na) * b  (32-bit limbs; carries run to the top limb)
template <int NS, class Acc>
__device__ __forceinline__ void mac_row(Acc &acc, const uint32_t *a, int na, uint32_t b, int off) {
  uint32_t carry = 0;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    if (k < off) continue;
    if (k - off < na) {
      const uint64_t t = (uint64_t)a[k - off] * b + (uint64_t)acc[k] + carry;
      acc[k] = (uint32_t)t;
      carry = (uint32_t)(t >> 32);
    } else {
      const uint64_t t = (uint64_t)acc[k] + carry;
      acc[k] = (uint32_t)t;
      carry = (uint32_t)(t >> 32);
    }
  }
}

// The lift of one coefficient with its limbs in `s` (WS limbs) and the
// product buffer `pr` (WS + 3 limbs), both zeroed; Acc is a register array
// or a strided view of scratch memory.
template <int WQ, class Acc>
__device__ __forceinline__ void crt_one(Acc &s, Acc &pr, const uint64_t *__restrict__ res, size_t stride,
                                        uint32_t L, const uint32_t *__restrict__ lift,
                                        const uint32_t *__restrict__ q, const uint32_t m[3],
                                        uint32_t shift, int wq) {
  constexpr int WS = WQ + 3;
  for (uint32_t j = 0; j < L; ++j) {
    const uint64_t r = res[(size_t)j * stride];
    const uint32_t *lj = lift + (size_t)j * wq;
    mac_row<WS>(s, lj, wq, (uint32_t)r, 0);
    mac_row<WS>(s, lj, wq, (uint32_t)(r >> 32), 1);
  }
  // pr = s * M
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    uint32_t carry = 0;
#pragma unroll
    for (int k = 0; k < WS + 3; ++k) {
      if (k < u) continue;
      const uint32_t sk = k - u < WS ? s[k - u] : 0u;
      const uint64_t t = (uint64_t)sk * m[u] + (uint64_t)pr[k] + carry;
      pr[k] = (uint32_t)t;
      carry = (uint32_t)(t >> 32);
    }
  }
  // tmp = floor(pr / 2^shift): 96 bits from limb shift/32 on (tmp < 2^75)
  const uint32_t sw = shift >> 5, sb = shift & 31;
  uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
#pragma unroll
  for (int k = 0; k < WS + 3; ++k) {
    const uint32_t v = pr[k];
    const uint32_t d = (uint32_t)k - sw;  // wraps for k < sw
    w0 = d == 0 ? v : w0;
    w1 = d == 1 ? v : w1;
    w2 = d == 2 ? v : w2;
    w3 = d == 3 ? v : w3;
  }
  uint32_t t3[3];
  t3[0] = sb ? (w0 >> sb) | (w1 << (32 - sb)) : w0;
  t3[1] = sb ? (w1 >> sb) | (w2 << (32 - sb)) : w1;
  t3[2] = sb ? (w2 >> sb) | (w3 << (32 - sb)) : w2;
  // pr = tmp * Q (WS limbs suffice: tmp * Q <= S)
#pragma unroll
  for (int k = 0; k < WS + 3; ++k) pr[k] = 0;
#pragma unroll
  for (int u = 0; u < 3; ++u) mac_row<WS>(pr, q, wq, t3[u], u);
  // s -= pr; then s -= Q if s >= Q
  uint32_t br = 0;
#pragma unroll
  for (int k = 0; k < WS; ++k) {
    const uint64_t t = (uint64_t)s[k] - pr[k] - br;
    s[k] = (uint32_t)t;
    br = (uint32_t)(t >> 63);
  }
  br = 0;
#pragma unroll
  for (int k = 0; k < WS; ++k) {
    const uint32_t qk = k < wq ? q[k] : 0u;
    const uint64_t t = (uint64_t)s[k] - qk - br;
    pr[k] = (uint32_t)t;
    br = (uint32_t)(t >> 63);
  }
  if (!br) {  // s >= Q
#pragma unroll
    for (int k = 0; k < WS; ++k) s[k] = pr[k];
  }
}

template <int WQ>
struct RegLimbs {
  uint32_t v[WQ + 6];
  __device__ __forceinline__ uint32_t &operator[](int k) { return v[k]; }
};

struct ScratchLimbs {
  uint32_t *base;
  size_t stride;
  __device__ __forceinline__ uint32_t &operator[](int k) { return base[(size_t)k * stride]; }
};

// The output is coefficient-major (coefficient i at words [i * limbs, (i+1) *
// limbs)), so a thread's own limbs are `limbs` consecutive words: stored
// directly, one store instruction would touch one cache line per lane.  The
// workgroup's coefficients are staged in LDS instead and written as one
// contiguous range, consecutive words across the lanes.
template <int WQ>
__global__ void __launch_bounds__(kCrtThreads)
k_crt_reg(uint64_t *__restrict__ out, const uint64_t *__restrict__ res, uint32_t n, uint32_t L,
          const uint32_t *__restrict__ lift, const uint32_t *__restrict__ q, uint32_t m0, uint32_t m1,
          uint32_t m2, uint32_t shift, uint32_t limbs) {
  __shared__ uint64_t stage[kCrtThreads * (WQ / 2)];
  const uint32_t c0 = blockIdx.x * kCrtThreads;
  const uint32_t i = c0 + threadIdx.x;
  if (i < n) {
    RegLimbs<WQ> s, pr;
#pragma unroll
    for (int k = 0; k < WQ + 6; ++k) s[k] = pr[k] = 0;
    const uint32_t m[3] = {m0, m1, m2};
    crt_one<WQ>(s, pr, res + i, n, L, lift, q, m, shift, WQ);
    uint64_t *o = stage + threadIdx.x * limbs;
#pragma unroll
    for (int t = 0; t < WQ / 2; ++t)
      if ((uint32_t)t < limbs) o[t] = (uint64_t)s[2 * t] | ((uint64_t)s[2 * t + 1] << 32);
  }
  __syncthreads();
  const uint32_t cnt = n - c0 < (uint32_t)kCrtThreads ? n - c0 : (uint32_t)kCrtThreads;
  const uint32_t words = cnt * limbs;
  uint64_t *dst = out + (size_t)c0 * limbs;
  for (uint32_t j = threadIdx.x; j < words; j += kCrtThreads) dst[j] = stage[j];
}

// Scratch path: s and pr of coefficient i at scratch[k * chunk + i] and
// scratch[(WS + k) * chunk + i]; coefficients [c0, c0 + cn) of the poly.
__global__ void __launch_bounds__(kCrtThreads)
k_crt_scratch(uint64_t *__restrict__ out, const uint64_t *__restrict__ res, uint32_t n, uint32_t c0,
              uint32_t cn, uint32_t L, const uint32_t *__restrict__ lift, const uint32_t *__restrict__ q,
              uint32_t m0, uint32_t m1, uint32_t m2, uint32_t shift, uint32_t limbs, int wq,
              uint32_t *__restrict__ scratch) {
  const uint32_t i = blockIdx.x * kCrtThreads + threadIdx.x;
  if (i >= cn) return;
  const int ws = wq + 3;
  ScratchLimbs s{scratch + i, cn}, pr{scratch + (size_t)ws * cn + i, cn};
  for (int k = 0; k < ws; ++k) s[k] = 0;
  for (int k = 0; k < ws + 3; ++k) pr[k] = 0;
  const uint32_t m[3] = {m0, m1, m2};
  const uint64_t *r = res + c0 + i;
  // runtime-width variant of crt_one
  for (uint32_t j = 0; j < L; ++j) {
    const uint64_t v = r[(size_t)j * n];
    const uint32_t *lj = lift + (size_t)j * wq;
    for (int h = 0; h < 2; ++h) {
      const uint32_t b = h ? (uint32_t)(v >> 32) : (uint32_t)v;
      uint32_t carry = 0;
      for (int k = h; k < ws; ++k) {
        const uint32_t a = k - h < wq ? lj[k - h] : 0u;
        const uint64_t t = (uint64_t)a * b + (uint64_t)s[k] + carry;
        s[k] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
      }
    }
  }
  for (int u = 0; u < 3; ++u) {
    uint32_t carry = 0;
    for (int k = u; k < ws + 3; ++k) {
      const uint32_t sk = k - u < ws ? s[k - u] : 0u;
      const uint64_t t = (uint64_t)sk * m[u] + (uint64_t)pr[k] + carry;
      pr[k] = (uint32_t)t;
      carry = (uint32_t)(t >> 32);
    }
  }
  const uint32_t sw = shift >> 5, sb = shift & 31;
  const uint32_t w0 = pr[sw], w1 = pr[sw + 1], w2 = pr[sw + 2];
  const uint32_t w3 = sw + 3 < (uint32_t)(ws + 3) ? pr[sw + 3] : 0u;
  uint32_t t3[3];
  t3[0] = sb ? (w0 >> sb) | (w1 << (32 - sb)) : w0;
  t3[1] = sb ? (w1 >> sb) | (w2 << (32 - sb)) : w1;
  t3[2] = sb ? (w2 >> sb) | (w3 << (32 - sb)) : w2;
  for (int k = 0; k < ws + 3; ++k) pr[k] = 0;
  for (int u = 0; u < 3; ++u) {
    uint32_t carry = 0;
    for (int k = u; k < ws; ++k) {
      const uint32_t a = k - u < wq ? q[k - u] : 0u;
      const uint64_t t = (uint64_t)a * t3[u] + (uint64_t)pr[k] + carry;
      pr[k] = (uint32_t)t;
      carry = (uint32_t)(t >> 32);
    }
  }
  uint32_t br = 0;
  for (int k = 0; k < ws; ++k) {
    const uint64_t t = (uint64_t)s[k] - pr[k] - br;
    s[k] = (uint32_t)t;
    br = (uint32_t)(t >> 63);
  }
  br = 0;
  for (int k = 0; k < ws; ++k) {
    const uint32_t qk = k < wq ? q[k] : 0u;
    const uint64_t t = (uint64_t)s[k] - qk - br;
    pr[k] = (uint32_t)t;
    br = (uint32_t)(t >> 63);
  }
  uint64_t *o = out + (size_t)(c0 + i) * limbs;
  for (uint32_t t = 0; t < limbs; ++t) {
    const uint32_t lo = br ? s[2 * t] : pr[2 * t];
    const uint32_t hi = br ? s[2 * t + 1] : pr[2 * t + 1];
    o[t] = (uint64_t)lo | ((uint64_t)hi << 32);
  }
}

namespace {

int crt_common(size_t L, const CrtHost **h) {
  if (L == 0 || L > ya_num_moduli()) return fail(YA_ERR_INVALID, "nmoduli %zu out of [1, %zu]", L, ya_num_moduli());
  *h = crt_host(L);
  return YA_OK;
}

}  // namespace

}  // namespace ya

using namespace ya;

extern "C" {

int ya_crt_limbs(size_t nmoduli, size_t *limbs) {
  if (!limbs) return fail(YA_ERR_INVALID, "null output");
  const CrtHost *h;
  int rc = crt_common(nmoduli, &h);
  if (rc) return rc;
  *limbs = h->limbs;
  return YA_OK;
}

int ya_crt_constants(size_t nmoduli, size_t limbs, uint64_t *moduli_product, uint64_t *modulus_shoup,
                     uint64_t *lifting, size_t *shift) {
  const CrtHost *h;
  int rc = crt_common(nmoduli, &h);
  if (rc) return rc;
  if (limbs != h->limbs) return fail(YA_ERR_INVALID, "limbs %zu != %zu for %zu moduli", limbs, h->limbs, nmoduli);
  if (moduli_product) std::memcpy(moduli_product, h->q.data(), limbs * 8);
  if (modulus_shoup) {
    modulus_shoup[0] = (uint64_t)h->m;
    modulus_shoup[1] = (uint64_t)(h->m >> 64);
  }
  if (lifting)
    for (size_t j = 0; j < nmoduli; ++j) std::memcpy(lifting + j * limbs, h->lift[j].data(), limbs * 8);
  if (shift) *shift = h->shift;
  return YA_OK;
}

int ya_poly_crt_lift(uint64_t *out, const uint64_t *residues, size_t degree, size_t nmoduli,
                     void *stream) {
  if (!out || !residues) return fail(YA_ERR_INVALID, "null pointer");
  if (degree == 0 || degree > ((size_t)1 << 20))
    return fail(YA_ERR_INVALID, "degree %zu out of [1, 2^20]", degree);
  const CrtHost *h;
  int rc = crt_common(nmoduli, &h);
  if (rc) return rc;
  const size_t in_end = degree * nmoduli, out_end = degree * h->limbs;
  if ((const uint64_t *)out < residues + in_end && residues < (const uint64_t *)out + out_end)
    return fail(YA_ERR_INVALID, "output must not overlap the residues");
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  const CrtDev *d;
  if ((rc = crt_device(nmoduli, h, &d))) return rc;
  hipStream_t s = (hipStream_t)stream;
  const uint32_t n = (uint32_t)degree, L = (uint32_t)nmoduli, sh = (uint32_t)h->shift,
                 lb = (uint32_t)h->limbs;
  hipError_t e = hipSuccess;
  if (nmoduli <= (size_t)kCrtMaxRegPrimes) {
    const dim3 grid((n + kCrtThreads - 1) / kCrtThreads), blk(kCrtThreads);
#define YA_CRT_CASE(W)                                                                          \
  case W:                                                                                       \
    hipLaunchKernelGGL(k_crt_reg<W>, grid, blk, 0, s, out, residues, n, L, d->lift, d->q, d->m[0], \
                       d->m[1], d->m[2], sh, lb);                                               \
    break;
    switch (d->wq) {
      YA_CRT_CASE(4)
      YA_CRT_CASE(8)
      YA_CRT_CASE(16)
      YA_CRT_CASE(24)
      YA_CRT_CASE(32)
      YA_CRT_CASE(48)
      YA_CRT_CASE(64)
      default: return fail(YA_ERR_INVALID, "crt limb bucket %d", d->wq);
    }
#undef YA_CRT_CASE
    e = hipGetLastError();
    return e == hipSuccess ? YA_OK : hip_fail(e, "crt lift launch");
  }
  const size_t chunk = std::min<size_t>(degree, kCrtChunk);
  const size_t words = (size_t)(2 * (d->wq + 3) + 3) * chunk;
  uint32_t *scratch = nullptr;
  e = hipMallocAsync((void **)&scratch, words * 4, s);
  if (e != hipSuccess) return fail(YA_ERR_ALLOC, "crt scratch: %s", hipGetErrorString(e));
  for (size_t c0 = 0; e == hipSuccess && c0 < degree; c0 += chunk) {
    const uint32_t cn = (uint32_t)std::min(chunk, degree - c0);
    hipLaunchKernelGGL(k_crt_scratch, dim3((cn + kCrtThreads - 1) / kCrtThreads), dim3(kCrtThreads), 0, s,
                       out, residues, n, (uint32_t)c0, cn, L, d->lift, d->q, d->m[0], d->m[1], d->m[2], sh,
                       lb, d->wq, scratch);
    e = hipGetLastError();
  }
  const hipError_t ef = hipFreeAsync(scratch, s);
  if (e == hipSuccess) e = ef;
  return e == hipSuccess ? YA_OK : hip_fail(e, "crt lift (scratch path)");
}

}  // extern "C"
