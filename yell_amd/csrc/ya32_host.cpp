// ya32_host.cpp -- host side of the 32-bit moduli build: constants, exact
// table construction, device-resident descriptors and the ya32_* C ABI.
//
// The reference selects its 32-bit build at compile time
// (CMakeLists.txt:19, YELL_USE_32BITS_MODULI -> include/yell/params_32.hpp):
// uint32 words, 291 primes 2^30 - k*2^16 + 1, degrees up to 2^15.  Here both
// builds live in one library: the 64-bit entry points are ya_*, these are
// ya32_*, each with its own tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/yell_amd.h"
#include "params32_table.inc"
#include "ya32.hpp"
#include "ya_internal.hpp"

namespace ya {
namespace {

const uint32_t kP32[YA32_NUM_MODULI] = YA32_TABLE_P;
const uint32_t kPn32[YA32_NUM_MODULI] = YA32_TABLE_PN;
const uint32_t kRoot32[YA32_NUM_MODULI] = YA32_TABLE_ROOT;

inline uint32_t mulm(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)((uint64_t)a * b % p); }

uint32_t powm(uint32_t a, uint64_t e, uint32_t p) {
  uint32_t r = 1 % p;
  while (e) {
    if (e & 1) r = mulm(r, a, p);
    a = mulm(a, a, p);
    e >>= 1;
  }
  return r;
}

// ops::shoupify (lib/ops.cpp:23-27) with 32-bit words
inline uint32_t shoup32(uint32_t w, uint32_t p) { return (uint32_t)(((uint64_t)w << 32) / p); }

inline uint32_t bitrev(uint32_t v, int bits) { return bits == 0 ? 0u : (__builtin_bitreverse32(v) >> (32 - bits)); }

struct Table32 {
  std::vector<Tw32> pairs;  // [fwd N | inv N]
  uint32_t sc[4];
};

// ntt<N>::ntt_precomputed::init (include/yell/details/ntt.hpp:27-77) with
// kMaxPolyDegree = 2^15: phi = root^(2^(15 - logn)); phi^i and phi^-i in
// bit-reversed order with Shoup companions; N^-1, N^-1 * invphi[1].
void build_table32(int logn, size_t cm, Table32 &t) {
  const size_t n = (size_t)1 << logn;
  const uint32_t p = kP32[cm];
  uint32_t phi = kRoot32[cm];
  for (int i = 0; i < kMaxLogN32 - logn; ++i) phi = mulm(phi, phi, p);
  std::vector<uint32_t> pw(n), ipw(n);
  uint32_t v = 1;
  for (size_t i = 0; i < n; ++i) {
    pw[i] = v;
    v = mulm(v, phi, p);
  }
  const uint32_t invphi = mulm(v, pw[n - 1], p);  // phi^N * phi^(N-1) = phi^-1
  v = 1;
  for (size_t i = 0; i < n; ++i) {
    ipw[i] = v;
    v = mulm(v, invphi, p);
  }
  t.pairs.resize(2 * n);
  for (size_t i = 0; i < n; ++i) {
    const size_t j = bitrev((uint32_t)i, logn);
    t.pairs[j] = Tw32{pw[i], shoup32(pw[i], p)};
    t.pairs[n + j] = Tw32{ipw[i], shoup32(ipw[i], p)};
  }
  t.sc[0] = powm((uint32_t)(n % p), p - 2, p);
  t.sc[2] = mulm(t.sc[0], t.pairs[n + 1].w, p);
  t.sc[1] = shoup32(t.sc[0], p);
  t.sc[3] = shoup32(t.sc[2], p);
}

struct Dev32 {
  ModDesc32 *d_desc = nullptr;  // [(kMaxLogN32 + 1) * YA32_NUM_MODULI]
  std::unique_ptr<std::atomic<int>[]> ready;
  std::vector<Tw32 *> tabs;
  std::mutex mu;
};

std::mutex g_mu32;
std::vector<std::unique_ptr<Dev32>> g_devs32;

inline size_t slot32(int logn, size_t cm) { return (size_t)logn * YA32_NUM_MODULI + cm; }

int dev32(Dev32 **out) {
  DeviceState *st;
  int rc = current_device(&st);  // makes the thread's device current
  if (rc) return rc;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  std::lock_guard<std::mutex> g(g_mu32);
  if ((int)g_devs32.size() <= dev) g_devs32.resize(dev + 1);
  if (!g_devs32[dev]) {
    auto d = std::make_unique<Dev32>();
    const size_t nslots = (size_t)(kMaxLogN32 + 1) * YA32_NUM_MODULI;
    e = hipMalloc(&d->d_desc, nslots * sizeof(ModDesc32));
    if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipMalloc(desc32): %s", hipGetErrorString(e));
    d->ready.reset(new std::atomic<int>[nslots]);
    for (size_t i = 0; i < nslots; ++i) d->ready[i].store(0);
    d->tabs.assign(nslots, nullptr);
    g_devs32[dev] = std::move(d);
  }
  *out = g_devs32[dev].get();
  return YA_OK;
}

int ensure32(Dev32 *dv, int logn, size_t cm0, size_t count) {
  bool all = true;
  for (size_t cm = cm0; cm < cm0 + count && all; ++cm)
    all = dv->ready[slot32(logn, cm)].load(std::memory_order_acquire) != 0;
  if (all) return YA_OK;
  std::lock_guard<std::mutex> g(dv->mu);
  const size_t n = (size_t)1 << logn;
  Table32 t;
  for (size_t cm = cm0; cm < cm0 + count; ++cm) {
    const size_t s = slot32(logn, cm);
    if (dv->ready[s].load(std::memory_order_acquire)) continue;
    build_table32(logn, cm, t);
    Tw32 *tab = nullptr;
    hipError_t e = hipMalloc(&tab, 2 * n * sizeof(Tw32));
    if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipMalloc(table32): %s", hipGetErrorString(e));
    e = hipMemcpy(tab, t.pairs.data(), 2 * n * sizeof(Tw32), hipMemcpyHostToDevice);
    ModDesc32 d{tab, tab + n, kP32[cm], kPn32[cm], t.sc[0], t.sc[1], t.sc[2], t.sc[3]};
    if (e == hipSuccess) e = hipMemcpy(dv->d_desc + s, &d, sizeof d, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(tab);
      return hip_fail(e, "upload table32");
    }
    dv->tabs[s] = tab;
    dv->ready[s].store(1, std::memory_order_release);
  }
  return YA_OK;
}

int shape32(size_t degree, size_t cm0, size_t nmoduli, int *logn) {
  if (degree < 2 || (degree & (degree - 1))) return fail(YA_ERR_INVALID, "degree %zu is not a power of two", degree);
  const int lg = __builtin_ctzll(degree);
  if (lg < kMinLogN32 || lg > kMaxLogN32)
    return fail(YA_ERR_INVALID, "degree 2^%d outside [2^%d, 2^%d] (32-bit moduli)", lg, kMinLogN32, kMaxLogN32);
  if (nmoduli == 0 || cm0 + nmoduli > YA32_NUM_MODULI)
    return fail(YA_ERR_INVALID, "moduli [%zu, %zu) outside [0, %d)", cm0, cm0 + nmoduli, YA32_NUM_MODULI);
  *logn = lg;
  return YA_OK;
}

int ntt32(uint32_t *x, size_t degree, size_t cm0, size_t nmod, size_t count, int dir, void *stream) {
  int logn;
  int rc = shape32(degree, cm0, nmod, &logn);
  if (rc) return rc;
  if (!x) return fail(YA_ERR_INVALID, "null buffer");
  if (count == 0 || count * nmod > (1u << 30)) return fail(YA_ERR_INVALID, "bad count %zu", count);
  Dev32 *dv;
  if ((rc = dev32(&dv)) || (rc = ensure32(dv, logn, cm0, nmod))) return rc;
  const ModDesc32 *d = dv->d_desc + slot32(logn, cm0);
  hipError_t e = dir == 2 ? launch32_backward(x, d, logn, nmod * count, nmod, (hipStream_t)stream)
                          : launch32_forward(x, d, logn, nmod * count, nmod, dir == 1, (hipStream_t)stream);
  return e == hipSuccess ? YA_OK : hip_fail(e, "NTT32 launch");
}

int pw32(PwOp op, uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmod, void *stream) {
  int logn;
  int rc = shape32(degree, 0, nmod, &logn);
  if (rc) return rc;
  if (!out || !a || (!b && op != PW_NEGATE)) return fail(YA_ERR_INVALID, "null buffer");
  Dev32 *dv;
  if ((rc = dev32(&dv)) || (rc = ensure32(dv, logn, 0, nmod))) return rc;
  hipError_t e = launch32_pointwise(op, out, a, b, dv->d_desc + slot32(logn, 0), logn, nmod, degree * nmod,
                                    (hipStream_t)stream);
  return e == hipSuccess ? YA_OK : hip_fail(e, "pointwise32 launch");
}

}  // namespace
}  // namespace ya

using namespace ya;

extern "C" {

size_t ya32_num_moduli(void) { return YA32_NUM_MODULI; }
uint32_t ya32_prime(size_t cm) { return cm < YA32_NUM_MODULI ? kP32[cm] : 0; }
uint32_t ya32_barrett_const(size_t cm) { return cm < YA32_NUM_MODULI ? kPn32[cm] : 0; }
uint32_t ya32_root(size_t cm) { return cm < YA32_NUM_MODULI ? kRoot32[cm] : 0; }

int ya32_ntt_table_host(size_t degree, size_t cm, uint32_t *phi, uint32_t *phi_shoup, uint32_t *invphi,
                        uint32_t *invphi_shoup, uint32_t *scalars) {
  int logn;
  int rc = shape32(degree, cm, 1, &logn);
  if (rc) return rc;
  if (!phi || !phi_shoup || !invphi || !invphi_shoup || !scalars) return fail(YA_ERR_INVALID, "null output");
  Table32 t;
  build_table32(logn, cm, t);
  for (size_t i = 0; i < degree; ++i) {
    phi[i] = t.pairs[i].w;
    phi_shoup[i] = t.pairs[i].ws;
    invphi[i] = t.pairs[degree + i].w;
    invphi_shoup[i] = t.pairs[degree + i].ws;
  }
  std::memcpy(scalars, t.sc, sizeof t.sc);
  return YA_OK;
}

int ya32_ntt_init_tables(size_t degree, size_t count) {
  int logn;
  int rc = shape32(degree, 0, count, &logn);
  if (rc) return rc;
  Dev32 *dv;
  if ((rc = dev32(&dv))) return rc;
  return ensure32(dv, logn, 0, count);
}

int ya32_ntt_forward(uint32_t *x, size_t degree, size_t cm, void *stream) { return ntt32(x, degree, cm, 1, 1, 0, stream); }
int ya32_ntt_forward_lazy(uint32_t *x, size_t degree, size_t cm, void *stream) {
  return ntt32(x, degree, cm, 1, 1, 1, stream);
}
int ya32_ntt_backward(uint32_t *x, size_t degree, size_t cm, void *stream) { return ntt32(x, degree, cm, 1, 1, 2, stream); }

int ya32_poly_forward(uint32_t *poly, size_t degree, size_t nmoduli, void *stream) {
  return ntt32(poly, degree, 0, nmoduli, 1, 0, stream);
}
int ya32_poly_forward_lazy(uint32_t *poly, size_t degree, size_t nmoduli, void *stream) {
  return ntt32(poly, degree, 0, nmoduli, 1, 1, stream);
}
int ya32_poly_backward(uint32_t *poly, size_t degree, size_t nmoduli, void *stream) {
  return ntt32(poly, degree, 0, nmoduli, 1, 2, stream);
}

int ya32_poly_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli, void *stream) {
  return pw32(PW_MUL, out, a, b, degree, nmoduli, stream);
}
int ya32_poly_add(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli, void *stream) {
  return pw32(PW_ADD, out, a, b, degree, nmoduli, stream);
}
int ya32_poly_sub(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli, void *stream) {
  return pw32(PW_SUB, out, a, b, degree, nmoduli, stream);
}
int ya32_poly_negate(uint32_t *x, size_t degree, size_t nmoduli, void *stream) {
  return pw32(PW_NEGATE, x, x, nullptr, degree, nmoduli, stream);
}
int ya32_poly_add_product_of(uint32_t *acc, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli,
                             void *stream) {
  return pw32(PW_MULADD, acc, a, b, degree, nmoduli, stream);
}

int ya32_poly_multiply_many(uint32_t *c, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli,
                            size_t count, void *stream) {
  int logn;
  int rc = shape32(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (!a || !b || !c) return fail(YA_ERR_INVALID, "null buffer");
  if (count == 0 || count * nmoduli > (1u << 30)) return fail(YA_ERR_INVALID, "bad count %zu", count);
  const size_t words = degree * nmoduli * count;
  if ((c < a + words && a < c + words) || (c < b + words && b < c + words))
    return fail(YA_ERR_INVALID, "c must not overlap a or b");
  Dev32 *dv;
  if ((rc = dev32(&dv)) || (rc = ensure32(dv, logn, 0, nmoduli))) return rc;
  hipError_t e = launch32_polymul(c, a, b, dv->d_desc + slot32(logn, 0), logn, nmoduli * count, nmoduli,
                                  (hipStream_t)stream);
  return e == hipSuccess ? YA_OK : hip_fail(e, "polymul32 launch");
}

int ya32_poly_multiply(uint32_t *c, const uint32_t *a, const uint32_t *b, size_t degree, size_t nmoduli,
                       void *stream) {
  return ya32_poly_multiply_many(c, a, b, degree, nmoduli, 1, stream);
}

}  // extern "C"
