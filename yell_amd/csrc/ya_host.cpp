// ya_host.cpp -- host runtime of the engine: constant tables, the
// device-resident NTT table cache, memory/stream plumbing and the C ABI of
// include/yell_amd.h for single polynomials.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/yell_amd.h"
#include "params_table.inc"
#include "ya_internal.hpp"
#include "ya_batch.hpp"
#include "ya_kernels.hpp"

namespace ya {

static const uint64_t kP[YA_NUM_MODULI] = YA_TABLE_P;
static const uint64_t kPn[YA_NUM_MODULI] = YA_TABLE_PN;
static const uint64_t kRoot[YA_NUM_MODULI] = YA_TABLE_ROOT;

static thread_local std::string t_err;
static thread_local int t_device = -1;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  t_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return fail(YA_ERR_LAUNCH, "%s: %s", what, hipGetErrorString(e));
}

uint64_t prime(size_t cm) { return kP[cm]; }

// ---------------------------------------------------------------------------
// Table construction (include/yell/details/ntt.hpp:27-77) with exact 128-bit
// arithmetic: every entry is a canonical residue or floor(w * 2^64 / p), so
// any exact evaluation reproduces the reference's words.
using u128 = unsigned __int128;

static inline uint64_t mulmod_exact(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((u128)a * b) % p);
}
static inline uint64_t shoupify(uint64_t w, uint64_t p) {
  return (uint64_t)(((u128)w << 64) / p);
}
static uint64_t powmod(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = mulmod_exact(r, a, p);
    a = mulmod_exact(a, a, p);
    e >>= 1;
  }
  return r;
}
static inline uint32_t bitrev(uint32_t v, int bits) {
  return bits == 0 ? 0u : (__builtin_bitreverse32(v) >> (32 - bits));
}

struct HostTable {
  std::vector<uint64_t> phi, phi_s, iphi, iphi_s;
  uint64_t sc[4];
};

static void build_table_host(int logn, size_t cm, HostTable &t) {
  const size_t n = (size_t)1 << logn;
  const uint64_t p = kP[cm];
  uint64_t phi = kRoot[cm];
  for (int i = 0; i < 20 - logn; ++i) phi = mulmod_exact(phi, phi, p);  // :39-43
  t.phi.resize(n);
  t.phi_s.resize(n);
  t.iphi.resize(n);
  t.iphi_s.resize(n);
  std::vector<uint64_t> pw(n), ipw(n);
  uint64_t v = 1;
  for (size_t i = 0; i < n; ++i) {  // :47-53
    pw[i] = v;
    v = mulmod_exact(v, phi, p);
  }
  const uint64_t invphi = mulmod_exact(v, pw[n - 1], p);  // :58, phi^N * phi^(N-1)
  v = 1;
  for (size_t i = 0; i < n; ++i) {  // :60-65
    ipw[i] = v;
    v = mulmod_exact(v, invphi, p);
  }
  for (size_t i = 0; i < n; ++i) {  // :68-71 (bit reversal)
    const size_t j = bitrev((uint32_t)i, logn);
    t.phi[j] = pw[i];
    t.iphi[j] = ipw[i];
  }
  for (size_t i = 0; i < n; ++i) {
    t.phi_s[i] = shoupify(t.phi[i], p);
    t.iphi_s[i] = shoupify(t.iphi[i], p);
  }
  t.sc[0] = powmod((uint64_t)n % p, p - 2, p);              // :73
  t.sc[2] = mulmod_exact(t.sc[0], n >= 2 ? t.iphi[1] : 1, p);  // :74
  t.sc[1] = shoupify(t.sc[0], p);
  t.sc[3] = shoupify(t.sc[2], p);
}

// ---------------------------------------------------------------------------
// Per-device state: the resident tables and descriptor array.
struct Slot {
  std::atomic<int> ready{0};
  Tw *dev = nullptr;           // [fwd N pairs | inv N pairs]
  uint64_t check[8] = {0};     // phi[1], phi_s[1], phi[N-1], phi_s[N-1], inv likewise
};

struct DeviceState {
  int device = 0;
  int cus = 256;
  hipStream_t tstream = nullptr;
  ModDesc *d_desc = nullptr;  // [(kMaxLogN + 1) * YA_NUM_MODULI]
  uint64_t *scratch = nullptr;  // this device's copy of ya::scratch (device address)
  std::mutex mu;
  std::unique_ptr<Slot[]> slots;  // same indexing as d_desc
};

// Scratch of ya_poly_multiply_noalloc: one static array per device, loaded
// with the code object, so that entry point never allocates.  Host code must
// not pass `scratch` itself to a kernel (on the host it names a stand-in, not
// device memory): DeviceState::scratch holds the device address that
// hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchWords = (size_t)1 << 21;  // 16 MiB
static __device__ __align__(256) uint64_t scratch[kScratchWords];

static std::mutex g_dev_mu;
static std::vector<std::unique_ptr<DeviceState>> g_devs;

static inline size_t slot_index(int logn, size_t cm) { return (size_t)logn * YA_NUM_MODULI + cm; }

int device_cus(DeviceState *st) { return st->cus; }

// 0: automatic, -1: never split, k >= 2: split every eligible launch by 2^k.
static std::atomic<int> g_split_mode{0};

int split_factor(DeviceState *st, int logn, size_t nres) {
  const int mode = g_split_mode.load(std::memory_order_relaxed);
  if (mode < 0 || logn < kSplitMinLog || logn > kSplitMaxLog) return 0;
  if (mode >= kSplitMinK)
    return (mode <= kSplitMaxK && logn - mode >= kSplitMinS) ? mode : 0;
  // automatic: when one workgroup per residue would leave at least half of
  // the CUs idle (its latency is one CU transforming a whole residue)
  if (nres * 2 > (size_t)st->cus) return 0;
  return split_choose_k(logn, nres, st->cus);
}

int current_device(DeviceState **out) {
  int dev = t_device;
  if (dev < 0) {
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(YA_ERR_NO_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    t_device = dev;
  }
  std::lock_guard<std::mutex> g(g_dev_mu);
  if ((int)g_devs.size() <= dev) g_devs.resize(dev + 1);
  if (!g_devs[dev]) {
    auto st = std::make_unique<DeviceState>();
    st->device = dev;
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return fail(YA_ERR_NO_DEVICE, "hipSetDevice(%d): %s", dev, hipGetErrorString(e));
    e = hipStreamCreateWithFlags(&st->tstream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(YA_ERR_NO_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    const size_t nslots = (size_t)(kMaxLogN + 1) * YA_NUM_MODULI;
    e = hipMalloc(&st->d_desc, nslots * sizeof(ModDesc));
    if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipMalloc(desc): %s", hipGetErrorString(e));
    e = hipMemset(st->d_desc, 0, nslots * sizeof(ModDesc));
    if (e != hipSuccess) return hip_fail(e, "hipMemset(desc)");
    st->slots.reset(new Slot[nslots]);
    // The stream-ordered allocations of the split and large-ring paths come
    // from the device's default pool; keeping freed blocks in the pool (no
    // release threshold) makes the next call's hipMallocAsync a reuse, not a
    // new mapping.
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
    e = hipGetSymbolAddress((void **)&st->scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !st->scratch)
      return fail(YA_ERR_NO_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      st->cus = cus;
    g_devs[dev] = std::move(st);
  }
  *out = g_devs[dev].get();
  return YA_OK;
}

// The calling thread's device state if ya_init (or an earlier call) created
// it; never creates it (that allocates).
static int existing_device(DeviceState **out) {
  int dev = t_device;
  if (dev < 0) {
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(YA_ERR_NO_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> g(g_dev_mu);
  if ((int)g_devs.size() <= dev || !g_devs[dev])
    return fail(YA_ERR_INVALID, "device %d is not initialised: call ya_init(%d) first", dev, dev);
  *out = g_devs[dev].get();
  return YA_OK;
}

// YA_OK if the tables of (2^logn, cm) for cm in [cm0, cm0 + count) are all
// resident (builds nothing).
static int tables_resident(DeviceState *st, int logn, size_t cm0, size_t count) {
  for (size_t cm = cm0; cm < cm0 + count; ++cm)
    if (!st->slots[slot_index(logn, cm)].ready.load(std::memory_order_acquire))
      return fail(YA_ERR_INVALID,
                  "tables of degree 2^%d, P[%zu] are not resident: call ya_ntt_init_tables(%zu, %zu) first", logn,
                  cm, (size_t)1 << logn, cm0 + count);
  return YA_OK;
}

// Builds (degree 2^logn, cm) for cm in [cm0, cm0 + count) where missing.
int ensure_tables(DeviceState *st, int logn, size_t cm0, size_t count) {
  if (logn < kMinLogN || logn > kMaxLogN) return fail(YA_ERR_INVALID, "unsupported log2(degree) %d", logn);
  if (cm0 + count > YA_NUM_MODULI || count == 0)
    return fail(YA_ERR_INVALID, "moduli [%zu, %zu) out of range", cm0, cm0 + count);
  bool all = true;
  for (size_t cm = cm0; cm < cm0 + count && all; ++cm)
    all = st->slots[slot_index(logn, cm)].ready.load(std::memory_order_acquire) != 0;
  if (all) return YA_OK;

  std::lock_guard<std::mutex> g(st->mu);
  std::vector<size_t> todo;
  for (size_t cm = cm0; cm < cm0 + count; ++cm)
    if (!st->slots[slot_index(logn, cm)].ready.load(std::memory_order_acquire)) todo.push_back(cm);
  if (todo.empty()) return YA_OK;

  const size_t n = (size_t)1 << logn;
  std::vector<HostTable> host(todo.size());
  {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const unsigned nth = (unsigned)std::min<size_t>(hw, todo.size());
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned k = 0; k < nth; ++k)
      pool.emplace_back([&]() {
        for (size_t j; (j = next.fetch_add(1)) < todo.size();) build_table_host(logn, todo[j], host[j]);
      });
    for (auto &th : pool) th.join();
  }
  hipError_t e = hipSetDevice(st->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  std::vector<Tw> pairs(2 * n);
  for (size_t j = 0; j < todo.size(); ++j) {
    const size_t cm = todo[j];
    const HostTable &t = host[j];
    Slot &s = st->slots[slot_index(logn, cm)];
    for (size_t i = 0; i < n; ++i) {
      pairs[i] = Tw{t.phi[i], t.phi_s[i]};
      pairs[n + i] = Tw{t.iphi[i], t.iphi_s[i]};
    }
    Tw *dev = nullptr;
    e = hipMalloc(&dev, 2 * n * sizeof(Tw));
    if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipMalloc(table 2^%d): %s", logn, hipGetErrorString(e));
    e = hipMemcpyAsync(dev, pairs.data(), 2 * n * sizeof(Tw), hipMemcpyHostToDevice, st->tstream);
    if (e != hipSuccess) return hip_fail(e, "upload table");
    ModDesc d;
    d.fwd = dev;
    d.inv = dev + n;
    d.p = kP[cm];
    d.pn = kPn[cm];
    d.inv_n = t.sc[0];
    d.inv_n_s = t.sc[1];
    d.inv_n_phi = t.sc[2];
    d.inv_n_phi_s = t.sc[3];
    e = hipMemcpyAsync(st->d_desc + slot_index(logn, cm), &d, sizeof d, hipMemcpyHostToDevice,
                       st->tstream);
    if (e != hipSuccess) return hip_fail(e, "upload descriptor");
    e = hipStreamSynchronize(st->tstream);  // `pairs` and `d` are reused / go out of scope
    if (e != hipSuccess) return hip_fail(e, "table stream sync");
    s.dev = dev;
    const uint64_t ck[8] = {t.phi[1], t.phi_s[1], t.phi[n - 1], t.phi_s[n - 1],
                            t.iphi[1], t.iphi_s[1], t.iphi[n - 1], t.iphi_s[n - 1]};
    std::memcpy(s.check, ck, sizeof ck);
    s.ready.store(1, std::memory_order_release);
  }
  return YA_OK;
}

const ModDesc *desc_ptr(DeviceState *st, int logn, size_t cm) {
  return st->d_desc + slot_index(logn, cm);
}

const ModDesc *desc_table_base(DeviceState *st) { return st->d_desc; }

int log2_degree(size_t degree) {
  if (degree < 2 || (degree & (degree - 1))) return -1;
  return __builtin_ctzll(degree);
}

int check_shape(size_t degree, size_t cm0, size_t nmoduli, int *logn) {
  const int l = log2_degree(degree);
  if (l < kMinLogN || l > kMaxLogN)
    return fail(YA_ERR_INVALID, "degree %zu unsupported (powers of two 2^%d..2^%d)", degree,
                kMinLogN, kMaxLogN);
  if (nmoduli == 0 || cm0 + nmoduli > YA_NUM_MODULI)
    return fail(YA_ERR_INVALID, "moduli [%zu, %zu) out of range [0, %d)", cm0, cm0 + nmoduli,
                YA_NUM_MODULI);
  *logn = l;
  return YA_OK;
}

// Pinned host allocations of this library (ya_host_alloc_pinned and the
// server's own buffers): base -> {bytes, device address of the base}.  A
// kernel may read or write host memory only inside one of these, so the
// served form's copy kernels take a request's buffers only after
// pinned_device_range found the whole range here.
struct PinnedBlock {
  size_t bytes;
  uintptr_t dev;
};
static std::mutex g_pin_mu;
static std::map<uintptr_t, PinnedBlock> g_pinned;

int pinned_alloc(void **ptr, size_t bytes) {
  *ptr = nullptr;
  hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipHostMalloc(%zu): %s", bytes, hipGetErrorString(e));
  void *dev = nullptr;
  if (hipHostGetDevicePointer(&dev, *ptr, 0) == hipSuccess && dev) {
    std::lock_guard<std::mutex> g(g_pin_mu);
    g_pinned[(uintptr_t)*ptr] = PinnedBlock{bytes, (uintptr_t)dev};
  } else {
    (void)hipGetLastError();  // usable for copies, not registered for kernels
  }
  return YA_OK;
}

int pinned_free(void *ptr) {
  if (!ptr) return YA_OK;
  {
    std::lock_guard<std::mutex> g(g_pin_mu);
    g_pinned.erase((uintptr_t)ptr);
  }
  const hipError_t e = hipHostFree(ptr);
  return e == hipSuccess ? YA_OK : hip_fail(e, "hipHostFree");
}

void *pinned_device_range(const void *p, size_t bytes) {
  if (!p) return nullptr;
  const uintptr_t a = (uintptr_t)p;
  std::lock_guard<std::mutex> g(g_pin_mu);
  auto it = g_pinned.upper_bound(a);
  if (it == g_pinned.begin()) return nullptr;
  --it;
  const uintptr_t base = it->first;
  if (a < base || bytes > it->second.bytes || a - base > it->second.bytes - bytes) return nullptr;
  return (void *)(it->second.dev + (a - base));
}

}  // namespace ya

using namespace ya;

// ===========================================================================
// C ABI
extern "C" {

const char *ya_status_string(int s) {
  switch (s) {
    case YA_OK: return "ok";
    case YA_ERR_INVALID: return "invalid argument";
    case YA_ERR_NO_DEVICE: return "no HIP device";
    case YA_ERR_ALLOC: return "allocation failed";
    case YA_ERR_LAUNCH: return "kernel launch / HIP call failed";
    case YA_ERR_UNSUPPORTED: return "unsupported";
    case YA_ERR_REJECTED: return "rejected by admission control";
    case YA_ERR_TIMEOUT: return "timed out";
    case YA_ERR_MISMATCH: return "tables do not match";
    default: return "unknown status";
  }
}

const char *ya_last_error(void) { return t_err.c_str(); }

int ya_init(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(YA_ERR_NO_DEVICE, "no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n) return fail(YA_ERR_INVALID, "device %d of %d", device, n);
  e = hipSetDevice(device);
  if (e != hipSuccess) return fail(YA_ERR_NO_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
  t_device = device;
  DeviceState *st;
  return current_device(&st);
}

int ya_device_count(int *count) {
  if (!count) return fail(YA_ERR_INVALID, "null count");
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return fail(YA_ERR_NO_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  return YA_OK;
}

int ya_device_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  return e == hipSuccess ? YA_OK : hip_fail(e, "hipDeviceSynchronize");
}

size_t ya_num_moduli(void) { return YA_NUM_MODULI; }
uint64_t ya_prime(size_t cm) { return cm < YA_NUM_MODULI ? kP[cm] : 0; }
uint64_t ya_barrett_const(size_t cm) { return cm < YA_NUM_MODULI ? kPn[cm] : 0; }
uint64_t ya_root(size_t cm) { return cm < YA_NUM_MODULI ? kRoot[cm] : 0; }

int ya_ntt_init_tables(size_t degree, size_t count) {
  int logn;
  int rc = check_shape(degree, 0, count, &logn);
  if (rc) return rc;
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  return ensure_tables(st, logn, 0, count);
}

int ya_ntt_table_host(size_t degree, size_t cm, uint64_t *phi, uint64_t *phi_s, uint64_t *iphi,
                      uint64_t *iphi_s, uint64_t *scalars) {
  const int logn = log2_degree(degree);
  if (logn < 1 || logn > 20 || cm >= YA_NUM_MODULI) return fail(YA_ERR_INVALID, "bad degree/cm");
  if (!phi || !phi_s || !iphi || !iphi_s || !scalars) return fail(YA_ERR_INVALID, "null output");
  HostTable t;
  build_table_host(logn, cm, t);
  std::memcpy(phi, t.phi.data(), degree * 8);
  std::memcpy(phi_s, t.phi_s.data(), degree * 8);
  std::memcpy(iphi, t.iphi.data(), degree * 8);
  std::memcpy(iphi_s, t.iphi_s.data(), degree * 8);
  std::memcpy(scalars, t.sc, sizeof t.sc);
  return YA_OK;
}

int ya_ntt_table_download(size_t degree, size_t cm, uint64_t *phi, uint64_t *phi_s, uint64_t *iphi,
                          uint64_t *iphi_s, uint64_t *scalars) {
  int logn;
  int rc = check_shape(degree, cm, 1, &logn);
  if (rc) return rc;
  if (!phi || !phi_s || !iphi || !iphi_s || !scalars) return fail(YA_ERR_INVALID, "null output");
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, cm, 1))) return rc;
  std::vector<Tw> pairs(2 * degree);
  ModDesc d;
  hipError_t e = hipMemcpy(pairs.data(), st->slots[slot_index(logn, cm)].dev,
                           2 * degree * sizeof(Tw), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "download table");
  e = hipMemcpy(&d, desc_ptr(st, logn, cm), sizeof d, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "download descriptor");
  for (size_t i = 0; i < degree; ++i) {
    phi[i] = pairs[i].w;
    phi_s[i] = pairs[i].ws;
    iphi[i] = pairs[degree + i].w;
    iphi_s[i] = pairs[degree + i].ws;
  }
  scalars[0] = d.inv_n;
  scalars[1] = d.inv_n_s;
  scalars[2] = d.inv_n_phi;
  scalars[3] = d.inv_n_phi_s;
  return YA_OK;
}

// ---- memory / streams / events -------------------------------------------
#define YA_HIP(expr, what)                         \
  do {                                             \
    hipError_t e_ = (expr);                        \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

int ya_malloc(void **ptr, size_t bytes) {
  if (!ptr) return fail(YA_ERR_INVALID, "null ptr");
  *ptr = nullptr;
  if (bytes == 0) return YA_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return YA_OK;
}
int ya_free(void *ptr) {
  if (ptr) YA_HIP(hipFree(ptr), "hipFree");
  return YA_OK;
}
int ya_memset_zero(void *ptr, size_t bytes, void *stream) {
  YA_HIP(hipMemsetAsync(ptr, 0, bytes, (hipStream_t)stream), "hipMemsetAsync");
  return YA_OK;
}
static int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind k, void *stream) {
  if (bytes == 0) return YA_OK;
  if (!dst || !src) return fail(YA_ERR_INVALID, "null pointer");
  YA_HIP(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream), "hipMemcpyAsync");
  YA_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return YA_OK;
}
int ya_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream) {
  return copy(dst, src, bytes, hipMemcpyHostToDevice, stream);
}
int ya_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream) {
  return copy(dst, src, bytes, hipMemcpyDeviceToHost, stream);
}
int ya_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream) {
  if (bytes == 0) return YA_OK;
  YA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream), "hipMemcpyAsync");
  return YA_OK;
}
int ya_host_alloc_pinned(void **ptr, size_t bytes) {
  if (!ptr) return fail(YA_ERR_INVALID, "null ptr");
  return pinned_alloc(ptr, bytes);
}
int ya_host_free_pinned(void *ptr) { return pinned_free(ptr); }
int ya_pointer_is_device(const void *ptr, int *is_device) {
  if (!is_device) return fail(YA_ERR_INVALID, "null output");
  *is_device = 0;
  if (!ptr) return YA_OK;
  hipPointerAttribute_t at;
  hipError_t e = hipPointerGetAttributes(&at, ptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory: not an error
    return YA_OK;
  }
  *is_device = at.type == hipMemoryTypeDevice ? 1 : 0;
  return YA_OK;
}

int ya_stream_create(void **stream) {
  if (!stream) return fail(YA_ERR_INVALID, "null stream");
  hipStream_t s;
  YA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  *stream = s;
  return YA_OK;
}
int ya_stream_destroy(void *stream) {
  if (stream) YA_HIP(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
  return YA_OK;
}
int ya_stream_synchronize(void *stream) {
  YA_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return YA_OK;
}
int ya_event_create(void **event) {
  if (!event) return fail(YA_ERR_INVALID, "null event");
  hipEvent_t ev;
  YA_HIP(hipEventCreate(&ev), "hipEventCreate");
  *event = ev;
  return YA_OK;
}
int ya_event_destroy(void *event) {
  if (event) YA_HIP(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
  return YA_OK;
}
int ya_event_record(void *event, void *stream) {
  YA_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
  return YA_OK;
}
int ya_event_synchronize(void *event) {
  YA_HIP(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
  return YA_OK;
}
int ya_event_elapsed_ms(float *ms, void *start, void *end) {
  if (!ms) return fail(YA_ERR_INVALID, "null ms");
  YA_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end), "hipEventElapsedTime");
  return YA_OK;
}

// ---- stream capture into a replayable graph --------------------------------
int ya_graph_begin(void *stream) {
  if (!stream) return fail(YA_ERR_INVALID, "capture needs a created stream, not the null stream");
  YA_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  return YA_OK;
}
int ya_graph_end(void *stream, void **graph) {
  if (!stream || !graph) return fail(YA_ERR_INVALID, "null stream or graph");
  *graph = nullptr;
  hipGraph_t g = nullptr;
  YA_HIP(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  const hipError_t ed = hipGraphDestroy(g);
  if (e == hipSuccess) e = ed;
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  *graph = x;
  return YA_OK;
}
int ya_graph_launch(void *graph, void *stream) {
  if (!graph) return fail(YA_ERR_INVALID, "null graph");
  YA_HIP(hipGraphLaunch((hipGraphExec_t)graph, (hipStream_t)stream), "hipGraphLaunch");
  return YA_OK;
}
int ya_graph_destroy(void *graph) {
  if (graph) YA_HIP(hipGraphExecDestroy((hipGraphExec_t)graph), "hipGraphExecDestroy");
  return YA_OK;
}

static bool overlaps(const uint64_t *a, size_t na, const uint64_t *b, size_t nb) {
  return a && b && a < b + nb && b < a + na;
}

// ---- NTT ------------------------------------------------------------------
// Paired forward transforms in the many-polynomial forms (on by default;
// ya_pair_forward(0) runs one residue per workgroup, for comparison).
static std::atomic<int> g_pair_forward{1};

static int ntt_common(uint64_t *x, size_t degree, size_t cm0, size_t nres, size_t count, int dir,
                      void *stream) {
  int logn;
  int rc = check_shape(degree, cm0, nres, &logn);
  if (rc) return rc;
  if (!x) return fail(YA_ERR_INVALID, "null buffer");
  if (count == 0 || count * nres > (1u << 30)) return fail(YA_ERR_INVALID, "bad count %zu", count);
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, cm0, nres))) return rc;
  const ModDesc *d = desc_ptr(st, logn, cm0);
  hipError_t e;
  const size_t total = nres * count;
  const NttShape sh = tuned_shape(logn, 0);
  if (dir == 2) {
    e = launch_backward(x, d, logn, total, nres, sh, (hipStream_t)stream);
  } else if (count >= 2 && g_pair_forward.load() && forward_pairs_ok(logn, sh)) {
    // same-prime residues of polynomials 2q, 2q+1 share a workgroup; an odd
    // last polynomial runs on its own
    const size_t np = count / 2;
    e = launch_forward_pairs(x, d, logn, np, nres, dir == 1, sh, (hipStream_t)stream);
    if (e == hipSuccess && (count & 1))
      e = launch_forward(x + 2 * np * nres * degree, d, logn, nres, nres, dir == 1, sh, (hipStream_t)stream);
  } else {
    e = launch_forward(x, d, logn, total, nres, dir == 1, sh, (hipStream_t)stream);
  }
  return e == hipSuccess ? YA_OK : hip_fail(e, "NTT launch");
}

int ya_ntt_forward(uint64_t *x, size_t degree, size_t cm, void *stream) {
  return ntt_common(x, degree, cm, 1, 1, 0, stream);
}
int ya_ntt_forward_lazy(uint64_t *x, size_t degree, size_t cm, void *stream) {
  return ntt_common(x, degree, cm, 1, 1, 1, stream);
}
int ya_ntt_backward(uint64_t *x, size_t degree, size_t cm, void *stream) {
  return ntt_common(x, degree, cm, 1, 1, 2, stream);
}
int ya_poly_forward(uint64_t *poly, size_t degree, size_t nmoduli, void *stream) {
  return ntt_common(poly, degree, 0, nmoduli, 1, 0, stream);
}
int ya_poly_forward_lazy(uint64_t *poly, size_t degree, size_t nmoduli, void *stream) {
  return ntt_common(poly, degree, 0, nmoduli, 1, 1, stream);
}
int ya_poly_backward(uint64_t *poly, size_t degree, size_t nmoduli, void *stream) {
  return ntt_common(poly, degree, 0, nmoduli, 1, 2, stream);
}
int ya_pair_forward(int on) {
  g_pair_forward.store(on ? 1 : 0);
  return YA_OK;
}
int ya_poly_forward_many(uint64_t *polys, size_t degree, size_t nmoduli, size_t count, void *stream) {
  return ntt_common(polys, degree, 0, nmoduli, count, 0, stream);
}
int ya_poly_forward_lazy_many(uint64_t *polys, size_t degree, size_t nmoduli, size_t count,
                              void *stream) {
  return ntt_common(polys, degree, 0, nmoduli, count, 1, stream);
}
int ya_poly_backward_many(uint64_t *polys, size_t degree, size_t nmoduli, size_t count, void *stream) {
  return ntt_common(polys, degree, 0, nmoduli, count, 2, stream);
}

int ya_poly_add_products(uint64_t *acc, const uint64_t *const *a, const uint64_t *const *b,
                         size_t npairs, size_t degree, size_t nmoduli, void *stream) {
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (!acc || (npairs && (!a || !b))) return fail(YA_ERR_INVALID, "null buffer");
  for (size_t j = 0; j < npairs; ++j)
    if (!a[j] || !b[j]) return fail(YA_ERR_INVALID, "null operand %zu", j);
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, 0, nmoduli))) return rc;
  for (size_t j0 = 0; j0 < npairs; j0 += kMacPairs) {
    MacPairs mp{};
    const int k = (int)std::min<size_t>(kMacPairs, npairs - j0);
    for (int j = 0; j < k; ++j) {
      mp.a[j] = a[j0 + j];
      mp.b[j] = b[j0 + j];
    }
    for (size_t r0 = 0; r0 < nmoduli; r0 += 65535) {
      const size_t nr = std::min<size_t>(65535, nmoduli - r0);
      MacPairs sub = mp;
      for (int j = 0; j < k; ++j) {
        sub.a[j] += r0 * degree;
        sub.b[j] += r0 * degree;
      }
      hipError_t e = launch_mac(acc + r0 * degree, sub, k, desc_ptr(st, logn, r0), degree, nr,
                                (hipStream_t)stream);
      if (e != hipSuccess) return hip_fail(e, "MAC launch");
    }
  }
  return YA_OK;
}

// acc = backward(acc + sum_j forward(a_j) * forward(b_j)) (accumulate), or
// backward(sum_j forward(a_j) * forward(b_j)) without reading acc; a_j, b_j
// are work space.
static int products_backward(uint64_t *acc, uint64_t *const *a, uint64_t *const *b, size_t npairs,
                             size_t degree, size_t nmoduli, bool accumulate, void *stream) {
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (!acc || (npairs && (!a || !b))) return fail(YA_ERR_INVALID, "null buffer");
  if (!accumulate && npairs == 0) return fail(YA_ERR_INVALID, "no pairs to sum");
  const size_t pw = degree * nmoduli;
  for (size_t j = 0; j < npairs; ++j) {
    if (!a[j] || !b[j]) return fail(YA_ERR_INVALID, "null operand %zu", j);
    if (overlaps(a[j], pw, acc, pw) || overlaps(b[j], pw, acc, pw) || overlaps(a[j], pw, b[j], pw))
      return fail(YA_ERR_INVALID, "operands of pair %zu overlap each other or acc", j);
    for (size_t k = 0; k < j; ++k)
      if (overlaps(a[j], pw, a[k], pw) || overlaps(a[j], pw, b[k], pw) || overlaps(b[j], pw, a[k], pw) ||
          overlaps(b[j], pw, b[k], pw))
        return fail(YA_ERR_INVALID, "operands of pairs %zu and %zu overlap", k, j);
  }
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, 0, nmoduli))) return rc;
  const ModDesc *d = desc_ptr(st, logn, 0);
  const hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (logn < kMinPairLog || nmoduli > 65535) {
    // small rings: the reference's sequence as separate launches (without an
    // accumulator to read, the first pair's product is written, not added)
    const NttShape sh = tuned_shape(logn, 0);
    for (size_t j = 0; e == hipSuccess && j < npairs; ++j) {
      e = launch_forward(a[j], d, logn, nmoduli, nmoduli, false, sh, s);
      if (e == hipSuccess) e = launch_forward(b[j], d, logn, nmoduli, nmoduli, false, sh, s);
    }
    size_t first = 0;
    if (e == hipSuccess && !accumulate) {
      e = launch_pointwise(PW_MUL, acc, a[0], b[0], d, degree, nmoduli, s);
      first = 1;
    }
    for (size_t j0 = first; e == hipSuccess && j0 < npairs; j0 += kMacPairs) {
      MacPairs mp{};
      const int k = (int)std::min<size_t>(kMacPairs, npairs - j0);
      for (int j = 0; j < k; ++j) {
        mp.a[j] = a[j0 + j];
        mp.b[j] = b[j0 + j];
      }
      for (size_t r0 = 0; e == hipSuccess && r0 < nmoduli; r0 += 65535) {
        const size_t nr = std::min<size_t>(65535, nmoduli - r0);
        MacPairs sub = mp;
        for (int j = 0; j < k; ++j) {
          sub.a[j] += r0 * degree;
          sub.b[j] += r0 * degree;
        }
        e = launch_mac(acc + r0 * degree, sub, k, d + r0, degree, nr, s);
      }
    }
    if (e == hipSuccess) e = launch_backward(acc, d, logn, nmoduli, nmoduli, sh, s);
    return e == hipSuccess ? YA_OK : hip_fail(e, "products_backward (separate launches)");
  }
  // chunks of kMacPairs pairs: each chunk's products go to its a_j; all but
  // the last chunk are added into acc (copied, for the first product when acc
  // is not read), the last one is summed by the inverse
  size_t j0 = 0;
  bool acc_live = accumulate;  // does acc hold words to add to?
  do {
    const int k = (int)std::min<size_t>(kMacPairs, npairs - j0);
    PairPtrs pp{};
    for (int j = 0; j < k; ++j) {
      pp.a[j] = a[j0 + j];
      pp.b[j] = b[j0 + j];
    }
    if (k > 0) e = launch_pair_products(pp, k, d, logn, nmoduli, tuned_mac_block_log(logn), s);
    j0 += k;
    if (e != hipSuccess) break;
    if (j0 < npairs) {
      for (int j = 0; e == hipSuccess && j < k; ++j) {
        if (!acc_live) {
          e = hipMemcpyAsync(acc, pp.a[j], pw * 8, hipMemcpyDeviceToDevice, s);
          acc_live = true;
        } else {
          e = launch_pointwise(PW_ADD, acc, acc, pp.a[j], d, degree, nmoduli, s);
        }
      }
    } else {
      e = launch_sum_backward(acc, pp, k, d, logn, nmoduli, acc_live, s);
    }
  } while (e == hipSuccess && j0 < npairs);
  return e == hipSuccess ? YA_OK : hip_fail(e, "products_backward");
}

int ya_poly_add_products_backward(uint64_t *acc, uint64_t *const *a, uint64_t *const *b, size_t npairs,
                                  size_t degree, size_t nmoduli, void *stream) {
  return products_backward(acc, a, b, npairs, degree, nmoduli, true, stream);
}

int ya_poly_sum_products_backward(uint64_t *out, uint64_t *const *a, uint64_t *const *b, size_t npairs,
                                  size_t degree, size_t nmoduli, void *stream) {
  return products_backward(out, a, b, npairs, degree, nmoduli, false, stream);
}

// Times the fresh-accumulator job (products_backward) on zero operands of
// the given shape with every accepted pair-pass block size -- the words do
// not change the timing -- after 0.15 s of untimed jobs (the clocks ramp up
// after an idle period), best of 5 taken round-robin, and keeps the fastest
// in the tune table (kind 4).
int ya_autotune_mac(size_t degree, size_t nmoduli, size_t npairs, double *report) {
  static const int kLogs[3] = {14, 13, 12};
  if (report)
    for (int k = 0; k < 3; ++k) report[k] = -1.0;
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (logn <= kMaxBlockLog) return fail(YA_ERR_INVALID, "degree 2^%d is one block: no pair-pass block size", logn);
  if (npairs == 0 || npairs > (size_t)kMacPairs || nmoduli > 65535)
    return fail(YA_ERR_INVALID, "tune MAC: %zu pairs (1..%d), %zu moduli (up to 65535)", npairs, kMacPairs,
                nmoduli);
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, 0, nmoduli))) return rc;
  const size_t pw = degree * nmoduli;
  uint64_t *buf = nullptr;
  hipError_t e = hipMalloc(&buf, (2 * npairs + 1) * pw * 8);
  if (e != hipSuccess) return fail(YA_ERR_ALLOC, "tune MAC buffers: %s", hipGetErrorString(e));
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  e = hipMemset(buf, 0, (2 * npairs + 1) * pw * 8);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  if (e != hipSuccess) rc = hip_fail(e, "tune MAC set-up");
  std::vector<uint64_t *> a(npairs), b(npairs);
  for (size_t j = 0; j < npairs; ++j) {
    a[j] = buf + 2 * j * pw;
    b[j] = buf + (2 * j + 1) * pw;
  }
  uint64_t *acc = buf + 2 * npairs * pw;
  const int saved = tuned_mac_block_log(logn);
  auto run = [&](int lb) {
    set_tuned_mac_block_log(logn, lb);
    return products_backward(acc, a.data(), b.data(), npairs, degree, nmoduli, false, s);
  };
  int cand[3], nc = 0;
  for (int lb : kLogs)
    if (lb == 14 || (lb == 13 && logn <= 18) || (lb == 12 && logn <= 17)) cand[nc++] = lb;
  float best[3] = {1e30f, 1e30f, 1e30f};
  if (rc == YA_OK) {
    const auto t0 = std::chrono::steady_clock::now();
    do {
      rc = run(cand[0]);
      if (rc == YA_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(YA_ERR_LAUNCH, "tune MAC settle");
    } while (rc == YA_OK && std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() < 0.15);
  }
  for (int i = 0; rc == YA_OK && i < nc; ++i) rc = run(cand[i]);
  for (int r = 0; rc == YA_OK && r < 5; ++r)
    for (int i = 0; rc == YA_OK && i < nc; ++i) {
      (void)hipEventRecord(ev0, s);
      rc = run(cand[i]);
      (void)hipEventRecord(ev1, s);
      if (hipEventSynchronize(ev1) != hipSuccess) rc = fail(YA_ERR_LAUNCH, "tune MAC sync");
      float ms = 0;
      (void)hipEventElapsedTime(&ms, ev0, ev1);
      if (rc == YA_OK && ms < best[i]) best[i] = ms;
    }
  int arg = saved;
  if (rc == YA_OK) {
    int bi = 0;
    for (int i = 1; i < nc; ++i)
      if (best[i] < best[bi]) bi = i;
    arg = cand[bi];
    if (report)
      for (int i = 0; i < nc; ++i) report[14 - cand[i]] = best[i];
  }
  set_tuned_mac_block_log(logn, arg);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(buf);
  return rc;
}

static int seam_host(uint64_t *x, size_t degree, const uint64_t *w, const uint64_t *ws, uint64_t p,
                     bool inverse) {
  int logn;
  int rc = check_shape(degree, 0, 1, &logn);
  if (rc) return rc;
  if (!x || !w || !ws) return fail(YA_ERR_INVALID, "null pointer");
  size_t cm = YA_NUM_MODULI;
  for (size_t k = 0; k < YA_NUM_MODULI; ++k)
    if (kP[k] == p) {
      cm = k;
      break;
    }
  if (cm == YA_NUM_MODULI) return fail(YA_ERR_INVALID, "modulus %llu is not in the prime table",
                                       (unsigned long long)p);
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, cm, 1))) return rc;
  const uint64_t *ck = st->slots[slot_index(logn, cm)].check + (inverse ? 4 : 0);
  if (w[1] != ck[0] || ws[1] != ck[1] || w[degree - 1] != ck[2] || ws[degree - 1] != ck[3])
    return fail(YA_ERR_MISMATCH, "caller tables are not the (degree %zu, P[%zu]) %s tables", degree,
                cm, inverse ? "inverse" : "forward");
  uint64_t *d = nullptr;
  hipError_t e = hipMalloc(&d, degree * 8);
  if (e != hipSuccess) return fail(YA_ERR_ALLOC, "hipMalloc: %s", hipGetErrorString(e));
  e = hipMemcpy(d, x, degree * 8, hipMemcpyHostToDevice);
  const ModDesc *dd = desc_ptr(st, logn, cm);
  if (e == hipSuccess) {
    if (inverse) e = launch_inverse_lazy(d, dd, logn, 1, nullptr);
    else e = launch_forward(d, dd, logn, 1, 1, true, tuned_shape(logn, 0), nullptr);
  }
  if (e == hipSuccess) e = hipMemcpy(x, d, degree * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? YA_OK : hip_fail(e, "seam");
}

int ya_negacylic_forward_lazy(uint64_t *x, size_t degree, const uint64_t *w, const uint64_t *ws,
                              uint64_t p) {
  return seam_host(x, degree, w, ws, p, false);
}
int ya_negacylic_backward_lazy(uint64_t *x, size_t degree, const uint64_t *w, const uint64_t *ws,
                               uint64_t p) {
  return seam_host(x, degree, w, ws, p, true);
}

// ---- pointwise ------------------------------------------------------------
static int pw_common(PwOp op, uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree,
                     size_t nmoduli, void *stream) {
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (!out || !a || (!b && op != PW_NEGATE)) return fail(YA_ERR_INVALID, "null buffer");
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, 0, nmoduli))) return rc;
  for (size_t r0 = 0; r0 < nmoduli; r0 += 65535) {
    const size_t nr = std::min<size_t>(65535, nmoduli - r0);
    const size_t off = r0 * degree;
    hipError_t e = launch_pointwise(op, out + off, a + off, b ? b + off : nullptr,
                                    desc_ptr(st, logn, r0), degree, nr, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "pointwise launch");
  }
  return YA_OK;
}

// The pointwise ops over `count` polynomials back to back; b_shared: one b
// for all.  A shared b must not overlap out (each polynomial's result would
// race with the others' reads of b).
static int pw_many(PwOp op, uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree,
                   size_t nmoduli, size_t count, int b_shared, void *stream) {
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (!out || !a || (!b && op != PW_NEGATE)) return fail(YA_ERR_INVALID, "null buffer");
  if (count == 0 || count * nmoduli > (1u << 30)) return fail(YA_ERR_INVALID, "bad count %zu", count);
  const size_t pw = degree * nmoduli;
  if (b && b_shared && b < out + pw * count && out < b + pw)
    return fail(YA_ERR_INVALID, "a shared b must not overlap out");
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, 0, nmoduli))) return rc;
  const size_t per = 65535 / nmoduli;
  for (size_t k0 = 0; k0 < count; k0 += per) {
    const size_t kc = std::min(per, count - k0);
    const uint64_t *bk = b ? (b_shared ? b : b + k0 * pw) : nullptr;
    hipError_t e = launch_pointwise_many(op, out + k0 * pw, a + k0 * pw, bk, desc_ptr(st, logn, 0), degree,
                                         nmoduli, kc, b_shared ? 0 : pw, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "pointwise launch");
  }
  return YA_OK;
}

int ya_poly_mul_many(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                     size_t count, int b_shared, void *stream) {
  return pw_many(PW_MUL, out, a, b, degree, nmoduli, count, b_shared, stream);
}
int ya_poly_add_many(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                     size_t count, int b_shared, void *stream) {
  return pw_many(PW_ADD, out, a, b, degree, nmoduli, count, b_shared, stream);
}
int ya_poly_sub_many(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                     size_t count, int b_shared, void *stream) {
  return pw_many(PW_SUB, out, a, b, degree, nmoduli, count, b_shared, stream);
}
int ya_poly_add_product_of_many(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t degree,
                                size_t nmoduli, size_t count, int b_shared, void *stream) {
  return pw_many(PW_MULADD, acc, a, b, degree, nmoduli, count, b_shared, stream);
}
int ya_poly_negate_many(uint64_t *x, size_t degree, size_t nmoduli, size_t count, void *stream) {
  return pw_many(PW_NEGATE, x, x, nullptr, degree, nmoduli, count, 0, stream);
}

// Element-wise op over polynomials [k0, k0 + kc) in launches of at most 65535
// residues; b_shared: b is one polynomial for all.
static hipError_t pw_range(PwOp op, uint64_t *out, const uint64_t *a, const uint64_t *b, const ModDesc *d,
                           size_t degree, size_t nmoduli, size_t kc, bool b_shared, hipStream_t s) {
  const size_t pw = degree * nmoduli, per = 65535 / nmoduli;
  for (size_t k0 = 0; k0 < kc; k0 += per) {
    const size_t n = std::min(per, kc - k0);
    const hipError_t e = launch_pointwise_many(op, out + k0 * pw, a + k0 * pw, b_shared ? b : b + k0 * pw, d,
                                               degree, nmoduli, n, b_shared ? 0 : pw, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}


int ya_poly_forward_lazy_add_product_many(uint64_t *acc, const uint64_t *add, const uint64_t *u,
                                          const uint64_t *key, size_t degree, size_t nmoduli, size_t count,
                                          void *stream) {
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (!acc || !u || !key) return fail(YA_ERR_INVALID, "null buffer");
  if (count == 0 || count * nmoduli > (1u << 30)) return fail(YA_ERR_INVALID, "bad count %zu", count);
  const size_t pw = degree * nmoduli, words = pw * count;
  if (overlaps(acc, words, add, words) || overlaps(acc, words, u, words) || overlaps(acc, words, key, pw))
    return fail(YA_ERR_INVALID, "add, u and key must not overlap acc");
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, 0, nmoduli))) return rc;
  const ModDesc *d = desc_ptr(st, logn, 0);
  const hipStream_t s = (hipStream_t)stream;
  const NttShape sh = tuned_shape(logn, 0);
  hipError_t e = hipSuccess;
  size_t done = 0;
  if (count >= 2 && g_pair_forward.load() && logn <= kMaxBlockLog && forward_pairs_ok(logn, sh)) {
    const size_t np = count / 2;
    e = launch_forward_pairs_mac(acc, add, u, key, d, logn, np, nmoduli, sh, s);
    done = 2 * np;
  }
  if (e == hipSuccess && done < count) {
    // the remaining polynomial(s): the same steps as separate launches
    const size_t kc = count - done;
    uint64_t *ak = acc + done * pw;
    if (add) e = pw_range(PW_ADD, ak, ak, add + done * pw, d, degree, nmoduli, kc, false, s);
    if (e == hipSuccess) e = launch_forward(ak, d, logn, kc * nmoduli, nmoduli, true, sh, s);
    if (e == hipSuccess) e = pw_range(PW_MULADD, ak, u + done * pw, key, d, degree, nmoduli, kc, true, s);
  }
  return e == hipSuccess ? YA_OK : hip_fail(e, "fused forward launch");
}

int ya_poly_add_product_backward_many(uint64_t *out, const uint64_t *acc, const uint64_t *x, const uint64_t *key,
                                      size_t degree, size_t nmoduli, size_t count, void *stream) {
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (!out || !acc || !x || !key) return fail(YA_ERR_INVALID, "null buffer");
  if (count == 0 || count * nmoduli > (1u << 30)) return fail(YA_ERR_INVALID, "bad count %zu", count);
  const size_t pw = degree * nmoduli, words = pw * count;
  if ((out != acc && overlaps(out, words, acc, words)) || overlaps(out, words, x, words) ||
      overlaps(out, words, key, pw))
    return fail(YA_ERR_INVALID, "out must be acc or not overlap acc, and must not overlap x or key");
  DeviceState *st;
  if ((rc = current_device(&st))) return rc;
  if ((rc = ensure_tables(st, logn, 0, nmoduli))) return rc;
  const ModDesc *d = desc_ptr(st, logn, 0);
  const hipStream_t s = (hipStream_t)stream;
  const NttShape sh = tuned_shape(logn, 0);
  hipError_t e;
  if (logn >= 3 && logn <= kMaxBlockLog) {
    e = launch_mac_inverse(out, acc, x, key, d, logn, nmoduli * count, nmoduli, sh, s);
  } else {
    e = out == acc ? hipSuccess : hipMemcpyAsync(out, acc, words * 8, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = pw_range(PW_MULADD, out, x, key, d, degree, nmoduli, count, true, s);
    if (e == hipSuccess) e = launch_backward(out, d, logn, nmoduli * count, nmoduli, sh, s);
  }
  return e == hipSuccess ? YA_OK : hip_fail(e, "fused backward launch");
}

int ya_poly_mul(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                void *stream) {
  return pw_common(PW_MUL, out, a, b, degree, nmoduli, stream);
}
int ya_poly_add(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                void *stream) {
  return pw_common(PW_ADD, out, a, b, degree, nmoduli, stream);
}
int ya_poly_sub(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                void *stream) {
  return pw_common(PW_SUB, out, a, b, degree, nmoduli, stream);
}
int ya_poly_negate(uint64_t *x, size_t degree, size_t nmoduli, void *stream) {
  return pw_common(PW_NEGATE, x, x, nullptr, degree, nmoduli, stream);
}
int ya_poly_add_product_of(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t degree,
                           size_t nmoduli, void *stream) {
  return pw_common(PW_MULADD, acc, a, b, degree, nmoduli, stream);
}

// noalloc (ya_poly_multiply_noalloc): the device state and the tables must
// already exist, and the scratch is the device's static array instead of a
// stream-ordered allocation.
static int polymul_impl(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree,
                        size_t nmoduli, size_t count, void *stream, bool noalloc) {
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (!a || !b || !c) return fail(YA_ERR_INVALID, "null buffer");
  if (count == 0 || count * nmoduli > (1u << 30)) return fail(YA_ERR_INVALID, "bad count %zu", count);
  const size_t words = degree * nmoduli * count;
  if ((c < a + words && a < c + words) || (c < b + words && b < c + words))
    return fail(YA_ERR_INVALID, "c must not overlap a or b");
  DeviceState *st;
  if (noalloc) {
    if ((rc = existing_device(&st))) return rc;
    if ((rc = tables_resident(st, logn, 0, nmoduli))) return rc;
  } else {
    if ((rc = current_device(&st))) return rc;
    if ((rc = ensure_tables(st, logn, 0, nmoduli))) return rc;
  }
  // scratch: the split launch keeps forward(b)'s top layers of every residue,
  // a large ring one polynomial's forward(b); the fused kernel needs none
  const size_t nres = nmoduli * count;
  const int ksplit = logn <= kMaxBlockLog ? split_factor(st, logn, nres) : 0;
  const size_t need = logn > kMaxBlockLog ? nmoduli << logn : ksplit ? nres << logn : 0;
  if (noalloc && need > kScratchWords)
    return fail(YA_ERR_INVALID,
                "degree %zu x %zu moduli needs %zu words of scratch; the static scratch holds %zu words", degree,
                nmoduli, need, kScratchWords);
  const ModDesc *d = desc_ptr(st, logn, 0);
  hipStream_t s = (hipStream_t)stream;
  uint64_t *tmp = nullptr;
  hipError_t e = hipSuccess;
  if (need) {
    if (noalloc) {
      tmp = st->scratch;
    } else if ((e = hipMallocAsync((void **)&tmp, need * 8, s)) != hipSuccess) {
      return fail(YA_ERR_ALLOC, "hipMallocAsync: %s", hipGetErrorString(e));
    }
  }
  if (logn <= kMaxBlockLog) {
    if (ksplit) {
      // few residues: spread each over 2^k workgroups; forward(b)'s top
      // layers go to the scratch (a and b are left untouched)
      const SplitArgs sa{nullptr, a, b, c, d, (uint32_t)nmoduli, tmp};
      e = launch_polymul_split(sa, logn, nres, ksplit, s);
    } else {
      e = launch_polymul(c, a, b, d, logn, nres, nmoduli, tuned_shape(logn, 1), s);
    }
  } else {
    // Large rings: the split launch -- the top forward layers of a (into c)
    // and b (into the scratch) per column, the rest of both forwards, the
    // product and the inverse's block layers per block into c, the top
    // inverse layers per column -- one polynomial's residues at a time
    const size_t pw = degree * nmoduli;
    const int k = large_ring_split_k(logn);
    for (size_t q = 0; e == hipSuccess && q < count; ++q) {
      const SplitArgs sa{nullptr, a + q * pw, b + q * pw, c + q * pw, d, (uint32_t)nmoduli, tmp};
      e = launch_polymul_split(sa, logn, nmoduli, k, s);
    }
  }
  if (tmp && !noalloc) {
    const hipError_t ef = hipFreeAsync(tmp, s);
    if (e == hipSuccess) e = ef;
  }
  return e == hipSuccess ? YA_OK : hip_fail(e, "polymul launch");
}

int ya_split_policy(int mode) {
  if (mode < -1 || mode == 1 || mode > kSplitMaxK) return fail(YA_ERR_INVALID, "split policy %d", mode);
  g_split_mode.store(mode, std::memory_order_relaxed);
  return YA_OK;
}

int ya_poly_multiply(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree,
                     size_t nmoduli, void *stream) {
  return polymul_impl(c, a, b, degree, nmoduli, 1, stream, false);
}

int ya_poly_multiply_noalloc(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree,
                             size_t nmoduli, void *stream) {
  return polymul_impl(c, a, b, degree, nmoduli, 1, stream, true);
}

int ya_poly_multiply_many(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree,
                          size_t nmoduli, size_t count, void *stream) {
  return polymul_impl(c, a, b, degree, nmoduli, count, stream, false);
}

}  // extern "C"
