// ya_server.cpp -- served form: a request loop with continuous batching.
//
// Clients submit poly-multiplication requests (host operands).  A dispatcher
// thread admits pending requests into a device arena bounded in bytes (the
// HBM budget): each request takes 3 * N * L words (a, b, c).  Whatever fits is
// batched into one launch of the persistent batched kernel on one of several
// in-flight slots (a stream with its own descriptor buffers), so requests
// arriving while earlier batches run are admitted as soon as memory and a
// slot are free, and their work overlaps on the device.  A completion thread
// retires batches as their events fire, copies results out, frees arena
// space and records each request's latency (submit -> result in the client's
// buffer).  Twiddle tables stay resident (the table cache).  The arithmetic is
// the batched kernel's, so every result equals ya_poly_multiply's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ya_batch.hpp"

namespace {

using clk = std::chrono::steady_clock;
using namespace ya;

struct Request {
  uint64_t id = 0;
  const uint64_t *a = nullptr, *b = nullptr;
  uint64_t *c = nullptr;
  uint32_t degree = 0, nmoduli = 0;
  size_t words = 0;   // N * L
  size_t off = 0;     // arena offset (words) of [a | b | c]
  clk::time_point t_submit, t_dispatch, t_done;
  int state = 0;      // 0 pending, 1 on device, 2 done, 3 failed
};

// First-fit allocator over [0, cap) words with coalescing.
class Arena {
 public:
  explicit Arena(size_t cap) : cap_(cap) { free_[0] = cap; }
  bool alloc(size_t n, size_t *off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used_ += n;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t cap() const { return cap_; }
  size_t used() const { return used_; }

 private:
  size_t cap_, used_ = 0;
  std::map<size_t, size_t> free_;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;  // batch start / kernel end (timing)
  clk::time_point host_dispatch;          // when the dispatcher took the batch
  double launch_us = 0;                   // host time spent issuing the batch
  BatchItem *h_items = nullptr;  // pinned host tables of the batch ...
  BatchRes *h_res = nullptr;
  const BatchItem *dh_items = nullptr;  // ... and their device addresses
  const BatchRes *dh_res = nullptr;
  BatchQueue *d_q = nullptr;
  uint64_t *d_scratch = nullptr;
  size_t cap_items = 0, cap_res = 0;
  // copy descriptors of the transfer kernels (pinned host memory, read by
  // the kernels through their device addresses): uploads, downloads
  XferDesc *h_xin = nullptr, *h_xout = nullptr;
  const XferDesc *d_xin = nullptr, *d_xout = nullptr;
  std::vector<std::shared_ptr<Request>> reqs;
  bool busy = false;
};

}  // namespace

struct ya_server {
  ya_server_config cfg{};
  DeviceState *st = nullptr;
  int device = 0;
  unsigned cus = 256;
  uint64_t *arena_base = nullptr;
  std::unique_ptr<Arena> arena;
  std::vector<Slot> slots;

  std::mutex mu;
  std::condition_variable cv_dispatch, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
  std::deque<int> inflight;  // slot indices in launch order
  bool stop = false;
  // bumped wherever a waiting thread is notified, so that spinning threads
  // can watch it without taking `mu`
  std::atomic<uint64_t> epoch{0};
  uint64_t next_id = 1;
  int error = YA_OK;

  // statistics
  std::vector<double> lat_us, queue_us, service_us;
  std::vector<double> b_gpu_us, b_d2h_us, b_launch_us, b_host_us;  // per batch
  uint64_t submitted = 0, completed = 0, rejected = 0, batches = 0, batched_jobs = 0;
  uint64_t max_pending = 0;

  std::thread th_dispatch, th_complete;
};

namespace {

void dispatch_loop(ya_server *s);
void complete_loop(ya_server *s);

int server_free(ya_server *s) {
  if (!s) return YA_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->epoch.fetch_add(1, std::memory_order_release);
  s->cv_dispatch.notify_all();
  s->cv_done.notify_all();
  if (s->th_dispatch.joinable()) s->th_dispatch.join();
  if (s->th_complete.joinable()) s->th_complete.join();
  {  // requests never dispatched fail, so that no waiter blocks forever
    std::lock_guard<std::mutex> g(s->mu);
    for (auto &r : s->pending) r->state = 3;
    s->pending.clear();
  }
  s->epoch.fetch_add(1, std::memory_order_release);
  s->cv_done.notify_all();
  (void)hipSetDevice(s->device);
  for (auto &sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.t0) (void)hipEventDestroy(sl.t0);
    if (sl.t1) (void)hipEventDestroy(sl.t1);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    (void)hipFree(sl.d_q);
    (void)hipFree(sl.d_scratch);
    (void)pinned_free(sl.h_items);
    (void)pinned_free(sl.h_res);
    (void)pinned_free(sl.h_xin);
    (void)pinned_free(sl.h_xout);
  }
  (void)hipFree(s->arena_base);
  delete s;
  return YA_OK;
}

// Launches one batch on slot `k` (called by the dispatcher without the lock).
int launch_slot(ya_server *s, int k) {
  Slot &sl = s->slots[k];
  std::vector<ya_polymul_job> jobs;
  jobs.reserve(sl.reqs.size());
  for (auto &r : sl.reqs) {
    uint64_t *base = s->arena_base + r->off;
    jobs.push_back(ya_polymul_job{base, base + r->words, base + 2 * r->words, r->degree, r->nmoduli});
  }
  // A lone request whose launch takes the split path addresses its residues
  // directly (its a, b, c are contiguous in the arena): no batch tables are
  // built or uploaded.
  int split_k = 0, lg = 0;
  bool direct = false;
  int rc;
  if (jobs.size() == 1) {
    lg = log2_degree(jobs[0].degree);
    if (lg >= kBatchMinLog && lg <= kBatchMaxLog) {
      if ((rc = ensure_tables(s->st, lg, 0, jobs[0].nmoduli))) return rc;
      split_k = split_factor(s->st, lg, jobs[0].nmoduli);
      direct = split_k != 0;
    }
  }
  std::vector<BatchItem> items;
  std::vector<BatchRes> res;
  bool has15 = false;
  BatchSegs segs{};
  hipError_t e = hipSuccess;
  if (!direct) {
    if ((rc = build_batch(s->st, jobs.data(), jobs.size(), items, res, &segs, &has15))) return rc;
    if (items.size() > sl.cap_items || res.size() > sl.cap_res)
      return fail(YA_ERR_INVALID, "batch exceeds slot capacity");
    if (has15 && !sl.d_scratch) e = hipMalloc(&sl.d_scratch, ((size_t)s->cus << 15) * 8);
    if (e != hipSuccess) return fail(YA_ERR_ALLOC, "slot scratch: %s", hipGetErrorString(e));
    bool one_class = true;
    for (const auto &it : items) one_class &= it.logn == items[0].logn;
    lg = (int)items[0].logn;
    split_k = one_class ? split_factor(s->st, lg, res.size()) : 0;
  }
  const size_t nres = direct ? jobs[0].nmoduli : res.size();
  // Uploads: a request whose a, b and c lie in pinned memory of this library
  // (16-byte aligned) is moved by the transfer kernel -- one launch for the
  // batch, no per-copy set-up -- the others by DMA copies.
  std::vector<uint8_t> by_kernel(jobs.size(), 0);
  size_t nin = 0, nout = 0, max_in = 0, max_out = 0;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const Request &r = *sl.reqs[j];
    const size_t bytes = r.words * 8;
    const auto aligned = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    const void *da = pinned_device_range(r.a, bytes), *db = pinned_device_range(r.b, bytes);
    void *dc = pinned_device_range(r.c, bytes);
    if (!da || !db || !dc || !aligned(da) || !aligned(db) || !aligned(dc) || (r.words & 1)) continue;
    if (nin + 2 > 65535 || nout + 1 > 65535) continue;  // one launch's descriptor limit
    by_kernel[j] = 1;
    if ((const uint64_t *)db == (const uint64_t *)da + r.words) {  // adjacent: b follows a in the arena too
      sl.h_xin[nin++] = XferDesc{(const uint64_t *)da, (uint64_t *)jobs[j].a, 2 * r.words};
      max_in = std::max(max_in, 2 * r.words);
    } else {
      sl.h_xin[nin++] = XferDesc{(const uint64_t *)da, (uint64_t *)jobs[j].a, r.words};
      sl.h_xin[nin++] = XferDesc{(const uint64_t *)db, (uint64_t *)jobs[j].b, r.words};
      max_in = std::max(max_in, r.words);
    }
    sl.h_xout[nout++] = XferDesc{jobs[j].c, (uint64_t *)dc, r.words};
    max_out = std::max(max_out, r.words);
  }
  if (e == hipSuccess) e = hipEventRecord(sl.t0, sl.stream);
  if (e == hipSuccess) e = launch_xfer(sl.d_xin, nin, max_in, sl.stream);
  for (size_t j = 0; j < jobs.size() && e == hipSuccess; ++j) {
    if (by_kernel[j]) continue;
    const Request &r = *sl.reqs[j];
    if (r.b == r.a + r.words) {  // a and b adjacent on the host too: one copy (b follows a in the arena)
      e = hipMemcpyAsync((void *)jobs[j].a, r.a, 2 * r.words * 8, hipMemcpyHostToDevice, sl.stream);
    } else {
      e = hipMemcpyAsync((void *)jobs[j].a, r.a, r.words * 8, hipMemcpyHostToDevice, sl.stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync((void *)jobs[j].b, r.b, r.words * 8, hipMemcpyHostToDevice, sl.stream);
    }
  }
  // The batch tables are written to the slot's pinned host arrays, which the
  // kernels read through their device addresses (a few KiB, read once per
  // item): no DMA copy, whose fixed set-up costs as much as a small batch's
  // transfer.
  if (!direct) {
    std::memcpy(sl.h_items, items.data(), items.size() * sizeof(BatchItem));
    std::memcpy(sl.h_res, res.data(), res.size() * sizeof(BatchRes));
  }
  if (e == hipSuccess) {
    if (split_k) {
      // few residues: spread each over 2^k workgroups; forward(b)'s top layers
      // go in place into the arena copy of b, which the server owns
      const SplitArgs sa = direct ? SplitArgs{nullptr, jobs[0].a, jobs[0].b, jobs[0].c, desc_ptr(s->st, lg, 0),
                                              (uint32_t)jobs[0].nmoduli, nullptr}
                                  : SplitArgs{sl.dh_res, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
      e = launch_polymul_split(sa, lg, nres, split_k, sl.stream);
    } else {
      const unsigned grid = (unsigned)std::min<size_t>(s->cus, items.size());
      e = launch_batch(sl.dh_items, segs, sl.dh_res, sl.d_q, sl.d_scratch, grid, sl.stream);
    }
  }
  if (e == hipSuccess) e = hipEventRecord(sl.t1, sl.stream);
  if (e == hipSuccess) e = launch_xfer(sl.d_xout, nout, max_out, sl.stream);
  for (size_t j = 0; j < jobs.size() && e == hipSuccess; ++j) {
    if (by_kernel[j]) continue;
    const Request &r = *sl.reqs[j];
    e = hipMemcpyAsync(r.c, jobs[j].c, r.words * 8, hipMemcpyDeviceToHost, sl.stream);
  }
  if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
  return e == hipSuccess ? YA_OK : hip_fail(e, "server launch");
}

// Threads that would block (the dispatcher, the completion thread, a
// waiting client) first poll for up to kSpinUs, yielding: waking a blocked
// thread costs tens of microseconds, as much as a small request's kernels.
constexpr int kSpinUs = 500;

template <class Pred>
void spin_then_wait(std::unique_lock<std::mutex> &lk, std::condition_variable &cv,
                    const std::atomic<uint64_t> &epoch, Pred pred) {
  const auto until = clk::now() + std::chrono::microseconds(kSpinUs);
  while (!pred()) {
    if (clk::now() >= until) {
      cv.wait(lk, pred);
      return;
    }
    const uint64_t seen = epoch.load(std::memory_order_acquire);
    lk.unlock();  // poll the epoch, not the mutex
    while (epoch.load(std::memory_order_acquire) == seen && clk::now() < until) std::this_thread::yield();
    lk.lock();
  }
}


void dispatch_loop(ya_server *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    int k = -1;
    spin_then_wait(lk, s->cv_dispatch, s->epoch, [&] {
      if (s->stop) return true;
      if (s->pending.empty()) return false;
      for (size_t i = 0; i < s->slots.size(); ++i)
        if (!s->slots[i].busy) return true;
      return false;
    });
    if (s->stop) return;
    for (size_t i = 0; i < s->slots.size(); ++i)
      if (!s->slots[i].busy) { k = (int)i; break; }
    Slot &sl = s->slots[k];
    sl.reqs.clear();
    size_t nres = 0;
    while (!s->pending.empty() && sl.reqs.size() < s->cfg.max_batch_jobs) {
      auto r = s->pending.front();
      if (nres + r->nmoduli > sl.cap_res) break;
      size_t off;
      if (!s->arena->alloc(3 * r->words, &off)) break;  // wait for memory to free up
      r->off = off;
      r->state = 1;
      r->t_dispatch = clk::now();
      nres += r->nmoduli;
      sl.reqs.push_back(r);
      s->pending.pop_front();
    }
    if (sl.reqs.empty()) {  // the head does not fit until a batch retires (which notifies)
      s->cv_dispatch.wait_for(lk, std::chrono::milliseconds(1));
      continue;
    }
    sl.busy = true;
    s->batches++;
    s->batched_jobs += sl.reqs.size();
    sl.host_dispatch = clk::now();
    lk.unlock();
    const int rc = launch_slot(s, k);
    const double lus = std::chrono::duration<double, std::micro>(clk::now() - sl.host_dispatch).count();
    // a launch that failed part-way may have queued copies into the arena and
    // the clients' buffers: they finish before that space is handed out again
    // and before the clients learn of the failure
    if (rc != YA_OK) (void)hipStreamSynchronize(sl.stream);
    lk.lock();
    sl.launch_us = lus;
    if (rc != YA_OK) {
      s->error = rc;
      for (auto &r : sl.reqs) {
        r->state = 3;
        s->arena->release(r->off, 3 * r->words);
      }
      sl.reqs.clear();
      sl.busy = false;
      s->epoch.fetch_add(1, std::memory_order_release);
      s->cv_done.notify_all();
      continue;
    }
    s->inflight.push_back(k);
    s->epoch.fetch_add(1, std::memory_order_release);
    s->cv_done.notify_all();  // wake the completion thread
  }
}

void complete_loop(ya_server *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    spin_then_wait(lk, s->cv_done, s->epoch, [&] { return s->stop || !s->inflight.empty(); });
    if (s->inflight.empty() && s->stop) return;
    std::vector<int> ready;
    std::vector<int> snapshot(s->inflight.begin(), s->inflight.end());
    lk.unlock();
    for (int k : snapshot) {
      hipError_t q = hipEventQuery(s->slots[k].done);
      if (q == hipSuccess) ready.push_back(k);
      else if (q != hipErrorNotReady) ready.push_back(-1 - k);  // failed
    }
    if (ready.empty()) std::this_thread::yield();  // a batch is in flight: poll its event again
    lk.lock();
    const auto now = clk::now();
    for (int code : ready) {
      const int k = code >= 0 ? code : -1 - code;
      Slot &sl = s->slots[k];
      if (code >= 0) {
        float g = 0, d2h = 0;
        if (hipEventElapsedTime(&g, sl.t0, sl.t1) == hipSuccess &&
            hipEventElapsedTime(&d2h, sl.t1, sl.done) == hipSuccess) {
          s->b_gpu_us.push_back(1e3 * g);
          s->b_d2h_us.push_back(1e3 * d2h);
        }
        s->b_launch_us.push_back(sl.launch_us);
        s->b_host_us.push_back(std::chrono::duration<double, std::micro>(now - sl.host_dispatch).count());
      }
      for (auto &r : sl.reqs) {
        r->t_done = now;
        r->state = code >= 0 ? 2 : 3;
        s->arena->release(r->off, 3 * r->words);
        if (code >= 0) {
          s->completed++;
          s->lat_us.push_back(std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count());
          s->queue_us.push_back(std::chrono::duration<double, std::micro>(r->t_dispatch - r->t_submit).count());
          s->service_us.push_back(std::chrono::duration<double, std::micro>(r->t_done - r->t_dispatch).count());
        }
      }
      sl.reqs.clear();
      sl.busy = false;
      s->inflight.erase(std::find(s->inflight.begin(), s->inflight.end(), k));
    }
    if (!ready.empty()) {
      s->epoch.fetch_add(1, std::memory_order_release);
      s->cv_dispatch.notify_all();
      s->cv_done.notify_all();
    }
  }
}

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int ya_server_create(ya_server **out, const ya_server_config *cfg_in) {
  if (!out) return fail(YA_ERR_INVALID, "null output");
  *out = nullptr;
  ya_server_config cfg{};
  cfg.hbm_budget_bytes = (size_t)1 << 30;
  cfg.max_batch_jobs = 256;
  cfg.max_inflight = 2;
  cfg.max_pending = 4096;
  cfg.max_moduli = 16;
  if (cfg_in) {
    if (cfg_in->hbm_budget_bytes) cfg.hbm_budget_bytes = cfg_in->hbm_budget_bytes;
    if (cfg_in->max_batch_jobs) cfg.max_batch_jobs = cfg_in->max_batch_jobs;
    if (cfg_in->max_inflight) cfg.max_inflight = cfg_in->max_inflight;
    if (cfg_in->max_pending) cfg.max_pending = cfg_in->max_pending;
    if (cfg_in->max_moduli) cfg.max_moduli = cfg_in->max_moduli;
  }
  if (cfg.max_inflight > 8 || cfg.max_batch_jobs > 65536 || cfg.max_moduli > 1000)
    return fail(YA_ERR_INVALID, "server config out of range");
  auto *s = new ya_server();
  s->cfg = cfg;
  int rc = current_device(&s->st);
  if (rc) {
    delete s;
    return rc;
  }
  (void)hipGetDevice(&s->device);
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) == hipSuccess && cus > 0)
    s->cus = (unsigned)cus;
  const size_t words = cfg.hbm_budget_bytes / 8;
  hipError_t e = hipMalloc(&s->arena_base, words * 8);
  if (e != hipSuccess) {
    delete s;
    return fail(YA_ERR_ALLOC, "arena of %zu bytes: %s", cfg.hbm_budget_bytes, hipGetErrorString(e));
  }
  s->arena.reset(new Arena(words));
  s->slots.resize(cfg.max_inflight);
  for (auto &sl : s->slots) {
    sl.cap_res = (size_t)cfg.max_batch_jobs * cfg.max_moduli;
    sl.cap_items = sl.cap_res;
    const BatchQueue q0{};
    e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&sl.done);
    if (e == hipSuccess) e = hipEventCreate(&sl.t0);
    if (e == hipSuccess) e = hipEventCreate(&sl.t1);
    if (e == hipSuccess) e = hipMalloc(&sl.d_q, sizeof(BatchQueue));
    if (e == hipSuccess) e = hipMemcpy(sl.d_q, &q0, sizeof q0, hipMemcpyHostToDevice);
    const size_t bi = sl.cap_items * sizeof(BatchItem), br = sl.cap_res * sizeof(BatchRes);
    if (e == hipSuccess &&
        (pinned_alloc((void **)&sl.h_items, bi) != YA_OK || pinned_alloc((void **)&sl.h_res, br) != YA_OK))
      e = hipErrorOutOfMemory;
    if (e == hipSuccess) {
      sl.dh_items = (const BatchItem *)pinned_device_range(sl.h_items, bi);
      sl.dh_res = (const BatchRes *)pinned_device_range(sl.h_res, br);
      if (!sl.dh_items || !sl.dh_res) e = hipErrorInvalidValue;
    }
    const size_t xin = 2 * (size_t)cfg.max_batch_jobs * sizeof(XferDesc), xout = xin / 2;
    if (e == hipSuccess && (pinned_alloc((void **)&sl.h_xin, xin) != YA_OK ||
                            pinned_alloc((void **)&sl.h_xout, xout) != YA_OK))
      e = hipErrorOutOfMemory;
    if (e == hipSuccess) {
      sl.d_xin = (const XferDesc *)pinned_device_range(sl.h_xin, xin);
      sl.d_xout = (const XferDesc *)pinned_device_range(sl.h_xout, xout);
      if (!sl.d_xin || !sl.d_xout) e = hipErrorInvalidValue;
    }
    if (e != hipSuccess) {
      server_free(s);
      return fail(YA_ERR_ALLOC, "server slot: %s", hipGetErrorString(e));
    }
  }
  s->th_dispatch = std::thread(dispatch_loop, s);
  s->th_complete = std::thread(complete_loop, s);
  *out = s;
  return YA_OK;
}

int ya_server_submit(ya_server *s, const uint64_t *a, const uint64_t *b, uint64_t *c,
                     uint32_t degree, uint32_t nmoduli, uint64_t *ticket) {
  if (!s || !a || !b || !c || !ticket) return fail(YA_ERR_INVALID, "null argument");
  int logn;
  int rc = check_shape(degree, 0, nmoduli, &logn);
  if (rc) return rc;
  if (logn < kBatchMinLog || logn > kBatchMaxLog)
    return fail(YA_ERR_UNSUPPORTED, "degree 2^%d outside the served range 2^%d..2^%d", logn,
                kBatchMinLog, kBatchMaxLog);
  const size_t words = (size_t)degree * nmoduli;
  auto r = std::make_shared<Request>();
  r->a = a;
  r->b = b;
  r->c = c;
  r->degree = degree;
  r->nmoduli = nmoduli;
  r->words = words;
  std::lock_guard<std::mutex> g(s->mu);
  if (s->stop) return fail(YA_ERR_REJECTED, "server stopped");
  if (nmoduli > s->cfg.max_moduli || 3 * words > s->arena->cap()) {
    s->rejected++;
    return fail(YA_ERR_REJECTED, "request of %zu bytes / %u moduli exceeds the server's bounds",
                3 * words * 8, nmoduli);
  }
  if (s->pending.size() >= s->cfg.max_pending) {
    s->rejected++;
    return fail(YA_ERR_REJECTED, "pending queue full (%u)", s->cfg.max_pending);
  }
  r->id = s->next_id++;
  r->t_submit = clk::now();
  s->pending.push_back(r);
  s->live[r->id] = r;
  s->submitted++;
  s->max_pending = std::max<uint64_t>(s->max_pending, s->pending.size());
  *ticket = r->id;
  s->epoch.fetch_add(1, std::memory_order_release);
  s->cv_dispatch.notify_one();
  return YA_OK;
}

int ya_server_wait(ya_server *s, uint64_t ticket, int timeout_ms, double *latency_us) {
  if (!s) return fail(YA_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return fail(YA_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
  auto r = it->second;
  auto pred = [&] { return r->state >= 2; };
  if (timeout_ms < 0) {
    spin_then_wait(lk, s->cv_done, s->epoch, pred);
  } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) {
    return fail(YA_ERR_TIMEOUT, "ticket %llu not done after %d ms", (unsigned long long)ticket, timeout_ms);
  }
  s->live.erase(r->id);
  if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
  if (r->state == 3) return fail(s->error ? s->error : YA_ERR_LAUNCH, "request failed on the device");
  return YA_OK;
}

int ya_server_poll(ya_server *s, uint64_t ticket, int *done) {
  if (!s || !done) return fail(YA_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->live.find(ticket);
  if (it == s->live.end()) return fail(YA_ERR_INVALID, "unknown ticket");
  *done = it->second->state >= 2;
  return YA_OK;
}

int ya_server_get_stats(ya_server *s, ya_server_stats *out) {
  if (!s || !out) return fail(YA_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  std::memset(out, 0, sizeof *out);
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->batches = s->batches;
  out->mean_batch_jobs = s->batches ? (double)s->batched_jobs / (double)s->batches : 0.0;
  out->max_pending = s->max_pending;
  out->arena_bytes_in_use = s->arena->used() * 8;
  out->p50_us = percentile(s->lat_us, 0.50);
  out->p90_us = percentile(s->lat_us, 0.90);
  out->p99_us = percentile(s->lat_us, 0.99);
  out->max_us = s->lat_us.empty() ? 0 : *std::max_element(s->lat_us.begin(), s->lat_us.end());
  double sum = 0;
  for (double v : s->lat_us) sum += v;
  out->mean_us = s->lat_us.empty() ? 0 : sum / (double)s->lat_us.size();
  out->p50_queue_us = percentile(s->queue_us, 0.50);
  out->p99_queue_us = percentile(s->queue_us, 0.99);
  out->p50_service_us = percentile(s->service_us, 0.50);
  out->p99_service_us = percentile(s->service_us, 0.99);
  out->p50_batch_gpu_us = percentile(s->b_gpu_us, 0.50);
  out->p99_batch_gpu_us = percentile(s->b_gpu_us, 0.99);
  out->p99_batch_d2h_us = percentile(s->b_d2h_us, 0.99);
  out->p99_batch_launch_us = percentile(s->b_launch_us, 0.99);
  out->p50_batch_host_us = percentile(s->b_host_us, 0.50);
  out->p99_batch_host_us = percentile(s->b_host_us, 0.99);
  return YA_OK;
}

int ya_server_reset_stats(ya_server *s) {
  if (!s) return fail(YA_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  s->lat_us.clear();
  s->queue_us.clear();
  s->service_us.clear();
  s->b_gpu_us.clear();
  s->b_d2h_us.clear();
  s->b_launch_us.clear();
  s->b_host_us.clear();
  s->submitted = s->completed = s->rejected = s->batches = s->batched_jobs = s->max_pending = 0;
  return YA_OK;
}

int ya_server_destroy(ya_server *s) { return server_free(s); }

// Poisson load generator: arrivals with exponential gaps at `rate_per_s` for
// `duration_s`; each request draws a shape uniformly from spec->shapes and an
// operand pair from a pool of `pool` pairs per shape (a uniform, b Gaussian
// sigma = 3.2, seeded).  With verify != 0 every served result is compared with
// the same pair multiplied alone (ya_poly_multiply, before the run).
int ya_server_run_poisson(ya_server *s, const ya_poisson_spec *spec, ya_serve_report *rep) {
  if (!s || !spec || !rep || spec->nshapes == 0 || !spec->degrees || !spec->nmoduli ||
      !(spec->rate_per_s > 0) || !(spec->duration_s > 0))
    return fail(YA_ERR_INVALID, "bad Poisson spec");
  const uint32_t pool = spec->pool ? spec->pool : 8;
  struct Shape {
    uint32_t n, L;
    size_t words;
    uint64_t *ops = nullptr;   // pinned: pool x [a | b]
    std::vector<uint64_t> want;  // pool x words (verify)
  };
  std::vector<Shape> shapes(spec->nshapes);
  int rc = YA_OK;
  auto cleanup = [&]() {
    for (auto &sh : shapes)
      if (sh.ops) (void)pinned_free(sh.ops);
  };
  for (uint32_t k = 0; k < spec->nshapes && rc == YA_OK; ++k) {
    Shape &sh = shapes[k];
    sh.n = spec->degrees[k];
    sh.L = spec->nmoduli[k];
    sh.words = (size_t)sh.n * sh.L;
    if (pinned_alloc((void **)&sh.ops, pool * 2 * sh.words * 8) != YA_OK) {
      rc = fail(YA_ERR_ALLOC, "pinned operand pool");
      break;
    }
    for (uint32_t q = 0; q < pool && rc == YA_OK; ++q) {
      uint64_t *a = sh.ops + (size_t)q * 2 * sh.words;
      rc = ya_sample_uniform(a, sh.n, sh.L, spec->seed + 1000003ull * k + 2 * q);
      if (rc == YA_OK) rc = ya_sample_gaussian(a + sh.words, sh.n, sh.L, 3.2, spec->seed + 1000003ull * k + 2 * q + 1);
    }
    if (rc == YA_OK && spec->verify) {  // the same pairs, each multiplied alone
      sh.want.resize((size_t)pool * sh.words);
      uint64_t *d = nullptr;
      if (hipMalloc(&d, 3 * sh.words * 8) != hipSuccess) { rc = fail(YA_ERR_ALLOC, "verify buffer"); break; }
      for (uint32_t q = 0; q < pool && rc == YA_OK; ++q) {
        const uint64_t *a = sh.ops + (size_t)q * 2 * sh.words;
        if (hipMemcpy(d, a, 2 * sh.words * 8, hipMemcpyHostToDevice) != hipSuccess) rc = fail(YA_ERR_LAUNCH, "verify copy");
        if (rc == YA_OK) rc = ya_poly_multiply(d + 2 * sh.words, d, d + sh.words, sh.n, sh.L, nullptr);
        if (rc == YA_OK && hipMemcpy(sh.want.data() + (size_t)q * sh.words, d + 2 * sh.words, sh.words * 8, hipMemcpyDeviceToHost) != hipSuccess)
          rc = fail(YA_ERR_LAUNCH, "verify copy back");
      }
      (void)hipFree(d);
    }
  }
  if (rc) {
    cleanup();
    return rc;
  }
  // Arrival schedule.
  std::mt19937_64 gen(spec->seed);
  std::exponential_distribution<double> gap(spec->rate_per_s);
  std::uniform_int_distribution<uint32_t> pick_shape(0, spec->nshapes - 1), pick_pair(0, pool - 1);
  struct Arr {
    double t;
    uint32_t shape, pair;
  };
  std::vector<Arr> arr;
  for (double t = gap(gen); t < spec->duration_s; t += gap(gen)) arr.push_back({t, pick_shape(gen), pick_pair(gen)});
  // Result buffers: a ring larger than the number of requests the server can
  // hold (pending + in flight); a waiter thread takes results in submission
  // order, verifies them and frees their ring entry.
  size_t max_words = 0;
  for (auto &sh : shapes) max_words = std::max(max_words, sh.words);
  size_t ring = 64;
  {
    std::lock_guard<std::mutex> g(s->mu);
    ring = std::max<size_t>(ring, (size_t)s->cfg.max_pending + (size_t)s->cfg.max_inflight * s->cfg.max_batch_jobs + 1);
  }
  ring = std::min(ring, arr.size() + 1);
  uint64_t *results = nullptr;
  if (pinned_alloc((void **)&results, ring * max_words * 8) != YA_OK) {
    cleanup();
    return fail(YA_ERR_ALLOC, "pinned result ring (%zu bytes)", ring * max_words * 8);
  }
  ya_server_reset_stats(s);
  std::vector<uint64_t> tickets(arr.size(), 0);
  std::vector<int> status(arr.size(), -1);  // -1 not yet submitted
  std::vector<uint8_t> retired(arr.size(), 0);
  std::mutex wmu;
  std::condition_variable wcv;
  size_t submitted_upto = 0, retired_upto = 0;  // ring entry i % ring is free once i < retired_upto + ring
  std::atomic<uint64_t> mismatches{0}, failures{0};
  uint64_t backpressure = 0;
  // kWaiters threads take results (waiter w: requests w, w + kWaiters, ...),
  // verify them and retire their ring entries; the ring advances over the
  // contiguous prefix of retired requests.  Verification reads two 1 MiB
  // buffers per request of 2^14 x 8 words: with 4 threads the verifiers fell behind near
  // 16 K req/s, the ring filled and arrivals were held back and released in
  // bursts (client_stalls), which the server saw as queueing -- p99 of 10-40
  // ms that a run without verification did not have (1.2 ms at 16 K req/s).
  // Half of the host's cores, 4 to 12, keep the checking off the path.
  const size_t kWaiters =
      std::max<size_t>(4, std::min<size_t>(12, std::thread::hardware_concurrency() / 2));
  std::vector<std::thread> waiters;
  for (size_t w = 0; w < kWaiters; ++w) {
    waiters.emplace_back([&, w]() {
      for (size_t i = w; i < arr.size(); i += kWaiters) {
        int st;
        {
          std::unique_lock<std::mutex> lk(wmu);
          wcv.wait(lk, [&] { return submitted_upto > i; });
          st = status[i];
        }
        if (st == YA_OK) {
          if (ya_server_wait(s, tickets[i], -1, nullptr) != YA_OK) {
            ++failures;
          } else if (spec->verify) {
            const Shape &sh = shapes[arr[i].shape];
            if (std::memcmp(results + (i % ring) * max_words, sh.want.data() + (size_t)arr[i].pair * sh.words,
                            sh.words * 8) != 0)
              ++mismatches;
          }
        }
        std::lock_guard<std::mutex> lk(wmu);
        retired[i] = 1;
        while (retired_upto < arr.size() && retired[retired_upto]) ++retired_upto;
        wcv.notify_all();
      }
    });
  }
  const auto t0 = clk::now();
  for (size_t i = 0; i < arr.size(); ++i) {
    const auto due = t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(arr[i].t));
    {
      std::unique_lock<std::mutex> lk(wmu);
      if (i >= retired_upto + ring) {
        ++backpressure;
        wcv.wait(lk, [&] { return i < retired_upto + ring; });
      }
    }
    while (clk::now() < due) {
      const auto left = due - clk::now();
      if (left > std::chrono::microseconds(200)) std::this_thread::sleep_for(left - std::chrono::microseconds(100));
    }
    const Shape &sh = shapes[arr[i].shape];
    const uint64_t *a = sh.ops + (size_t)arr[i].pair * 2 * sh.words;
    const int st = ya_server_submit(s, a, a + sh.words, results + (i % ring) * max_words, sh.n, sh.L, &tickets[i]);
    std::lock_guard<std::mutex> lk(wmu);
    status[i] = st;
    submitted_upto = i + 1;
    wcv.notify_all();
  }
  for (auto &t : waiters) t.join();
  const double wall = std::chrono::duration<double>(clk::now() - t0).count();
  ya_server_stats st;
  ya_server_get_stats(s, &st);
  std::memset(rep, 0, sizeof *rep);
  rep->offered = arr.size();
  rep->completed = st.completed;
  rep->rejected = st.rejected;
  rep->failed = failures;
  rep->mismatches = mismatches;
  rep->wall_s = wall;
  rep->throughput_per_s = wall > 0 ? (double)st.completed / wall : 0;
  rep->p50_us = st.p50_us;
  rep->p90_us = st.p90_us;
  rep->p99_us = st.p99_us;
  rep->max_us = st.max_us;
  rep->mean_us = st.mean_us;
  rep->batches = st.batches;
  rep->mean_batch_jobs = st.mean_batch_jobs;
  rep->max_pending = st.max_pending;
  rep->client_stalls = backpressure;
  rep->p50_queue_us = st.p50_queue_us;
  rep->p99_queue_us = st.p99_queue_us;
  rep->p50_service_us = st.p50_service_us;
  rep->p99_service_us = st.p99_service_us;
  rep->p50_batch_gpu_us = st.p50_batch_gpu_us;
  rep->p99_batch_gpu_us = st.p99_batch_gpu_us;
  rep->p99_batch_d2h_us = st.p99_batch_d2h_us;
  rep->p99_batch_launch_us = st.p99_batch_launch_us;
  rep->p50_batch_host_us = st.p50_batch_host_us;
  rep->p99_batch_host_us = st.p99_batch_host_us;
  (void)pinned_free(results);
  cleanup();
  return YA_OK;
}

}  // extern "C"
