// ya_batch.hip -- batched poly-multiplication over jobs of uneven degree and
// modulus count: one persistent kernel drains a queue of work items.
//
// A work item is one degree class (2^LB words per residue) and up to
// 2^(10+LE-LB) residues (for LB <= 14) that one 1024-thread workgroup
// transforms side by side: each residue gets a sub-group of 2^(LB-LE)
// threads and its own slice of the workgroup's LDS, so a 2^10 residue and a
// 2^14 residue keep every lane busy alike.  A 2^15 residue is one item: its
// top layer runs in registers as the words arrive, its two 2^14 halves
// through LDS one after the other (the upper half's inputs parked in global
// memory meanwhile).
// Arithmetic and layer order are those of the plain kernels (the reference's
// butterflies, lib/ntt.cpp:47-83), so each result equals ya_poly_multiply's.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/yell_amd.h"
#include "ya_batch.hpp"
#include "ya_blocks.hpp"
#include "ya_internal.hpp"

namespace ya {

// Barriers of the sub-groups (batch_small): a sub-group of one wave or less
// needs no barrier between its passes (WaveSync, ya_blocks.hpp).

// The fused poly-multiplication of one sub-group's residue (sub-group `sub`
// of 2^(LB-LE) threads), its exchanges separated by `sync`.
template <int LB, int LE, class S>
__device__ __forceinline__ void batch_small_body(const BatchItem &it, const BatchRes *__restrict__ res,
                                                 uint64_t *lds, uint32_t sub, uint32_t t, S &&sync) {
  constexpr uint32_t TS = 1u << (LB - LE);  // threads per residue
  const uint32_t tl = t % TS;
  const bool active = sub < it.count;
  BatchRes r;
  ModDesc d;
  if constexpr (LB - LE >= 6) {
    // a sub-group spans whole waves: its residue is wave-uniform, so keep the
    // pointers and constants in scalar registers
    const uint32_t k = __builtin_amdgcn_readfirstlane(it.first + (active ? sub : 0u));
    r = res[k];
    d = load_desc(r.d);
  } else {
    r = res[it.first + (active ? sub : 0u)];
    d = *r.d;
  }
  uint64_t *my_lds = lds + sub * LdsWords<LB>::value;
  uint64_t x[1 << LE], y[1 << LE];
  if constexpr (LB >= 10) {
    // both forward transforms at once (shared twiddles, no parking in c); a
    // sub-group without a residue transforms the first one and stores nothing
    constexpr int W0 = FirstWindow<LB, LE>::value;
    gmem_load<W0, LE>(r.a, x, tl);
    gmem_load<W0, LE>(r.b, y, tl);
    fwd_block_dual<LB, LE>(x, y, my_lds, (const TwC *)d.fwd, LB, 0u, d.p, tl, sync);
  } else {
    fwd_block<LB, LE>(x, r.a, my_lds, (const TwC *)d.fwd, LB, 0u, d.p, tl, sync);
    if (active) gmem_store<0, LE>(r.c, x, tl);
    const uint32_t t2 = phase_v(tl);
    fwd_block<LB, LE>(x, r.b, my_lds, (const TwC *)d.fwd + phase_zero_s(), LB, 0u, d.p, t2, sync);
    if (active) {
      gmem_load<0, LE>(r.c, y, t2);
    } else {  // no residue: any defined words will do, nothing is stored
#pragma unroll
      for (int e = 0; e < (1 << LE); ++e) y[e] = x[e];
    }
  }
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(y[e], x[e], d.p, d.pn);
  const uint32_t t3 = phase_v(tl);
  inv_block<LB, LE, TOP_MERGED>(x, active ? r.c : nullptr, my_lds, (const TwC *)d.inv, LB, 0u, d, t3, sync);
}

// Sub-grouped fused poly-multiplication of up to 2^(10+LE-LB) residues.
// A sub-group of whole waves orders its exchanges that cross its waves with
// the workgroup's barrier: every sub-group of an item runs the same body (one
// without a residue computes on the item's first and stores nothing), so all
// of them reach the same barriers; the exchanges inside each wave need the
// wave only (kWaveLocal).  (A barrier of each sub-group's own, in LDS, paid
// while every exchange had a barrier; with only the crossing exchanges left
// the workgroup's measured as fast or faster: 2^13 1.086 -> 1.067-1.072 ms on
// the tuner's batch.)
template <int LB, int LE>
__device__ __forceinline__ void batch_small(const BatchItem &it, const BatchRes *__restrict__ res,
                                            uint64_t *lds, uint32_t t) {
  constexpr uint32_t TS = 1u << (LB - LE);  // threads per residue
  // a sub-group of whole waves has a wave-uniform index: take it from lane
  // 0's thread index (a scalar), so the LDS slice address is scalar too
  const uint32_t sub = TS >= (uint32_t)kBatchThreads ? 0u
                       : TS >= 64u        ? __builtin_amdgcn_readfirstlane(t) / TS
                                          : t / TS;
  if constexpr (TS > 64u) {
    batch_small_body<LB, LE>(it, res, lds, sub, t, WgSync());
  } else {
    batch_small_body<LB, LE>(it, res, lds, sub, t, WaveSync());
  }
}


// Bit 14 of one operand of a 2^15 poly-multiplication in registers: thread t
// loads its first-window words t | e << 10 of the lower half and their
// partners 2^14 above, keeps the lower outputs in x and parks the upper ones
// at the same offsets in `park` (in two groups of eight, to bound registers).
__device__ __forceinline__ void top15_fwd_keep(const uint64_t *__restrict__ src, uint64_t (&x)[16],
                                               uint64_t *__restrict__ park, const Tw &w, const PrimeK &k,
                                               uint32_t t) {
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    uint64_t up[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const uint32_t i = t | ((uint32_t)(g * 8 + e) << 10);
      x[g * 8 + e] = src[i];
      up[e] = src[i + (1u << 14)];
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      ct_butterfly(x[g * 8 + e], up[e], w, k);
      park[t | ((uint32_t)(g * 8 + e) << 10)] = up[e];
    }
  }
}

// One 2^15 residue; `scr` is this workgroup's private 2^14-word scratch.
// Bit 14 of both operands in registers (the upper outputs parked: a's in
// c's upper half, b's in scr); the lower half's remaining forward layers of
// both at once (shared twiddles), product and inverse layers below bit 14,
// written to c's lower half; then the upper half from the parked words, its
// inverse outputs kept in registers for the merged top inverse layer with
// the lower half's words.  Every word parked in global memory is read back
// by the thread that wrote it, at the same address (the first window of a
// 2^14 block and the last window of its inverse are both t | e << 10).
__device__ __forceinline__ void batch_15(const BatchItem &it, const BatchRes *__restrict__ res,
                                         uint64_t *lds, uint64_t *__restrict__ scr, uint32_t t) {
  constexpr int LE = 4;
  constexpr int W0 = FirstWindow<14, LE>::value;
  static_assert(W0 == 10, "first window of a 2^14 block");
  const BatchRes r = res[it.first];
  const ModDesc d = load_desc(r.d);
  uint64_t x[1 << LE], y[1 << LE];
  {
    const Tw w = load_tw(d.fwd + 1);  // (N >> 15) + 0
    const PrimeK k(d.p);
    top15_fwd_keep(r.a, x, r.c + (1u << 14), w, k, t);
    top15_fwd_keep(r.b, y, scr, w, k, t);
  }
  fwd_block_dual<14, LE>(x, y, lds, (const TwC *)d.fwd + phase_zero_s(), 15, 0u, d.p, t);
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(y[e], x[e], d.p, d.pn);
  const uint32_t t1 = phase_v(t);
  inv_block<14, LE, TOP_GS>(x, r.c, lds, (const TwC *)d.inv + phase_zero_s(), 15, 0u, d, t1);
  const uint32_t t2 = phase_v(t);
  gmem_load<W0, LE>(r.c + (1u << 14), x, t2);
  gmem_load<W0, LE>(scr, y, t2);
  // the lower half's inverse may still be reading the LDS in other waves
  fwd_block_dual<14, LE, true>(x, y, lds, (const TwC *)d.fwd + phase_zero_s(), 15, 1u << 14, d.p, t2);
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) x[e] = mulmod_barrett(y[e], x[e], d.p, d.pn);
  const uint32_t t3 = phase_v(t);
  inv_block<14, LE, TOP_GS>(x, nullptr, lds, (const TwC *)d.inv + phase_zero_s(), 15, 1u << 14, d, t3);
  // merged top layer (include/yell/details/ntt.hpp:151-159) on words i, i + 2^14
  const uint64_t p = d.p, p2 = p << 1;
#pragma unroll
  for (int e = 0; e < (1 << LE); ++e) {
    const uint32_t i = t3 | ((uint32_t)e << W0);
    const uint64_t u = r.c[i], v = x[e];
    r.c[i] = mulmod_shoup(u + v, d.inv_n, d.inv_n_s, p);
    r.c[i + (1u << 14)] = mulmod_shoup(u - v + p2, d.inv_n_phi, d.inv_n_phi_s, p);
  }
}

// The lane number (0..63), computed where it is written: the asm is volatile
// so the value is not hoisted out of the item loop and kept in a register.
__device__ __forceinline__ uint32_t lane_id_opaque() {
  uint32_t lane;
  __asm__ volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(lane));
  return lane;
}

// Thread 0 claims the next work item: its own XCD segment first, then the
// others in turn; the segment cursor lives in LDS (no per-thread register is
// carried across items).  0xffffffff when every segment is drained.
__device__ __forceinline__ uint32_t claim_item(const BatchSegs &segs, BatchQueue *q, uint32_t &seg,
                                               uint32_t &tried) {
  while (tried < kSegs) {
    const uint32_t n = segs.begin[seg + 1] - segs.begin[seg];
    if (n != 0) {
      const uint32_t j = atomicAdd(&q->next[seg], 1u);
      if (j < n) return segs.begin[seg] + j;
    }
    seg = (seg + 1) % kSegs;
    ++tried;
  }
  return 0xffffffffu;
}

__global__ void __launch_bounds__(kBatchThreads)
k_batch_polymul(const BatchItem *__restrict__ items, BatchSegs segs,
                const BatchRes *__restrict__ res, BatchQueue *q, uint64_t *__restrict__ scratch,
                uint32_t stagger) {
  __shared__ uint64_t lds[LdsWords<14>::value];
  stagger_start(stagger);
  // claim-ahead: item i is read from s_item[i & 1] while thread 0 claims
  // item i + 1 into the other slot, so the queue atomic's round trip overlaps
  // the work and one barrier per item remains
  __shared__ uint32_t s_item[2], s_seg, s_tried;
  if (threadIdx.x == 0) {
    s_seg = blockIdx.x % kSegs;
    s_tried = 0;
    s_item[0] = claim_item(segs, q, s_seg, s_tried);
  }
  __syncthreads();
  // The thread index is rebuilt per item from the wave's base (a scalar) and
  // the lane number, so no vector register holds it across the items (one
  // that did was spilled to scratch, the bodies needing all 128).
  const uint32_t wave_base = __builtin_amdgcn_readfirstlane(threadIdx.x & ~63u);
  for (uint32_t i = 0;; ++i) {
    const uint32_t k = s_item[i & 1u];
    if (k == 0xffffffffu) break;
    // per item, so that nothing derived from the thread index is hoisted out
    // of the loop (and kept live across every item)
    const uint32_t t = wave_base | lane_id_opaque();
    if (t == 0) s_item[(i + 1) & 1u] = claim_item(segs, q, s_seg, s_tried);
    const BatchItem it = items[k];
    switch (it.logn * 8 + it.le) {
      case 7 * 8 + 3: batch_small<7, 3>(it, res, lds, t); break;
      case 7 * 8 + 4: batch_small<7, 4>(it, res, lds, t); break;
      case 8 * 8 + 3: batch_small<8, 3>(it, res, lds, t); break;
      case 8 * 8 + 4: batch_small<8, 4>(it, res, lds, t); break;
      case 9 * 8 + 3: batch_small<9, 3>(it, res, lds, t); break;
      case 9 * 8 + 4: batch_small<9, 4>(it, res, lds, t); break;
      case 10 * 8 + 3: batch_small<10, 3>(it, res, lds, t); break;
      case 10 * 8 + 4: batch_small<10, 4>(it, res, lds, t); break;
      case 11 * 8 + 3: batch_small<11, 3>(it, res, lds, t); break;
      case 11 * 8 + 4: batch_small<11, 4>(it, res, lds, t); break;
      case 12 * 8 + 3: batch_small<12, 3>(it, res, lds, t); break;
      case 12 * 8 + 4: batch_small<12, 4>(it, res, lds, t); break;
      case 13 * 8 + 3: batch_small<13, 3>(it, res, lds, t); break;
      case 13 * 8 + 4: batch_small<13, 4>(it, res, lds, t); break;
      case 14 * 8 + 4: batch_small<14, 4>(it, res, lds, t); break;
      case 15 * 8 + 4: batch_15(it, res, lds, scratch + ((size_t)blockIdx.x << 14), t); break;
      default: break;  // rejected by the host
    }
    // the item's LDS is dead before the next item's first store, and the
    // next claim is visible
    __syncthreads();
  }
  // The last workgroup to leave resets the queue for the next launch.
  if ((wave_base | lane_id_opaque()) == 0) {
    __threadfence();
    if (atomicAdd(&q->done, 1u) == gridDim.x - 1) {
      for (int i = 0; i < kSegs; ++i) q->next[i] = 0;
      q->done = 0;
      __threadfence();
    }
  }
}

}  // namespace ya

// ===========================================================================
using namespace ya;

struct ya_batch_plan {
  int device = 0;
  size_t nitems = 0, nres = 0;
  BatchSegs segs{};
  unsigned grid = 0;
  BatchItem *d_items = nullptr;
  BatchRes *d_res = nullptr;
  BatchQueue *d_q = nullptr;
  uint64_t *d_scratch = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double total_polymuls = 0;
  int form = YA_BATCH_PERSISTENT;  // launch form chosen at creation
  int grid_logn = 0;               // degree class of the grid and split forms
  int split_k = 0;                 // split form: 2^k workgroups per residue
  uint64_t *d_split = nullptr;     // split form: scratch [nres][2^grid_logn]
  // grid form over back-to-back polynomials of one shape: the fused kernel
  // addresses residues from its launch index (no residue-table load)
  bool strided = false;
  const uint64_t *sa = nullptr, *sb = nullptr;
  uint64_t *sc = nullptr;
  size_t s_nmod = 0;
  const ModDesc *s_desc = nullptr;
};

namespace {

std::atomic<int> g_form_override{0};
std::atomic<int> g_strided{1};  // ya_batch_strided(0): always the residue table (tests, comparison)

int plan_free(ya_batch_plan *pl) {
  if (!pl) return YA_OK;
  if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  (void)hipFree(pl->d_items);
  (void)hipFree(pl->d_res);
  (void)hipFree(pl->d_q);
  (void)hipFree(pl->d_scratch);
  (void)hipFree(pl->d_split);
  delete pl;
  return YA_OK;
}

}  // namespace

namespace ya {

int build_batch(DeviceState *st, const ya_polymul_job *jobs, size_t njobs,
                std::vector<BatchItem> &items, std::vector<BatchRes> &res, BatchSegs *segs,
                bool *has15) {
  std::vector<std::vector<BatchRes>> by_class(kBatchMaxLog + 1);
  *has15 = false;
  int rc;
  for (size_t j = 0; j < njobs; ++j) {
    const ya_polymul_job &jb = jobs[j];
    int logn;
    if ((rc = check_shape(jb.degree, 0, jb.nmoduli, &logn))) return rc;
    if (logn < kBatchMinLog || logn > kBatchMaxLog)
      return fail(YA_ERR_UNSUPPORTED, "job %zu: degree 2^%d outside the batched range 2^%d..2^%d",
                  j, logn, kBatchMinLog, kBatchMaxLog);
    if (!jb.a || !jb.b || !jb.c) return fail(YA_ERR_INVALID, "job %zu: null buffer", j);
    const size_t words = (size_t)jb.degree * jb.nmoduli;
    if ((jb.c < jb.a + words && jb.a < jb.c + words) || (jb.c < jb.b + words && jb.b < jb.c + words))
      return fail(YA_ERR_INVALID, "job %zu: c overlaps a or b", j);
    if ((rc = ensure_tables(st, logn, 0, jb.nmoduli))) return rc;
    for (uint32_t r = 0; r < jb.nmoduli; ++r) {
      const size_t off = (size_t)r * jb.degree;
      by_class[logn].push_back(BatchRes{jb.a + off, jb.b + off, jb.c + off, desc_ptr(st, logn, r)});
    }
    *has15 |= logn == 15;
  }
  // Items per (degree, prime) class; whole classes are dealt to the XCD
  // segments by estimated work (N log N per residue), largest first to the
  // least loaded segment; each segment is then ordered largest degree first.
  struct Cls {
    int lg;
    const ModDesc *d;
    std::vector<BatchRes> v;
    double work;
  };
  std::vector<Cls> classes;
  for (int lg = kBatchMaxLog; lg >= kBatchMinLog; --lg) {
    std::map<const ModDesc *, std::vector<BatchRes>> byd;
    for (const auto &r : by_class[lg]) byd[r.d].push_back(r);
    for (auto &kv : byd)
      classes.push_back(Cls{lg, kv.first, std::move(kv.second), 0.0});
  }
  for (auto &c : classes) c.work = (double)c.v.size() * (double)(1u << c.lg) * c.lg;
  std::stable_sort(classes.begin(), classes.end(), [](const Cls &x, const Cls &y) { return x.work > y.work; });
  double load[kSegs] = {0};
  std::vector<std::vector<const Cls *>> seg_cls(kSegs);
  for (const auto &c : classes) {
    int best = 0;
    for (int k = 1; k < kSegs; ++k)
      if (load[k] < load[best]) best = k;
    load[best] += c.work;
    seg_cls[best].push_back(&c);
  }
  items.clear();
  res.clear();
  for (int sgi = 0; sgi < kSegs; ++sgi) {
    segs->begin[sgi] = (uint32_t)items.size();
    auto &lst = seg_cls[sgi];
    std::stable_sort(lst.begin(), lst.end(), [](const Cls *x, const Cls *y) { return x->lg > y->lg; });
    for (const Cls *c : lst) {
      const int lg = c->lg;
      int le = lg >= 14 ? 4 : tuned_shape(lg, 2).le;
      le = (le == 3) ? 3 : 4;
      const uint32_t per = lg >= 14 ? 1u : (1u << (10 + le - lg));
      for (size_t i = 0; i < c->v.size(); i += per) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(per, c->v.size() - i);
        items.push_back(BatchItem{(uint32_t)lg, (uint32_t)le, (uint32_t)res.size(), cnt});
        res.insert(res.end(), c->v.begin() + i, c->v.begin() + i + cnt);
      }
    }
  }
  segs->begin[kSegs] = (uint32_t)items.size();
  return YA_OK;
}

hipError_t launch_batch(const BatchItem *items, const BatchSegs &segs, const BatchRes *res,
                        BatchQueue *q, uint64_t *scratch, unsigned grid, hipStream_t s) {
  // every workgroup is resident (one per CU): the whole grid is the first
  // wave, staggered when each workgroup has at least 8 items to run
  const int steps = tuned_stagger(kBatchMaxLog);
  const uint32_t stagger = (steps > 0 && segs.begin[kSegs] >= 8u * grid && grid <= 0xffffffu)
                               ? (grid << 8) | (uint32_t)steps
                               : 0u;
  hipLaunchKernelGGL(k_batch_polymul, dim3(grid), dim3(kBatchThreads), 0, s, items, segs, res, q,
                     scratch, stagger);
  return hipGetLastError();
}

}  // namespace ya

extern "C" {

int ya_batch_plan_create(ya_batch_plan **out, const ya_polymul_job *jobs, size_t njobs) {
  if (!out || (!jobs && njobs)) return fail(YA_ERR_INVALID, "null argument");
  *out = nullptr;
  if (njobs == 0) return fail(YA_ERR_INVALID, "empty batch");
  DeviceState *st;
  int rc = current_device(&st);
  if (rc) return rc;
  std::vector<BatchItem> items;
  std::vector<BatchRes> res;
  bool has15 = false;
  BatchSegs segs{};
  if ((rc = build_batch(st, jobs, njobs, items, res, &segs, &has15))) return rc;
  auto *pl = new ya_batch_plan();
  pl->device = 0;
  (void)hipGetDevice(&pl->device);
  pl->nitems = items.size();
  pl->segs = segs;
  pl->nres = res.size();
  int cus = 0;
  hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device);
  if (e != hipSuccess || cus <= 0) cus = 256;
  pl->grid = (unsigned)std::min<size_t>((size_t)cus, items.size());
  e = hipMalloc(&pl->d_items, items.size() * sizeof(BatchItem));
  if (e == hipSuccess) e = hipMalloc(&pl->d_res, res.size() * sizeof(BatchRes));
  if (e == hipSuccess) e = hipMalloc(&pl->d_q, sizeof(BatchQueue));
  if (e == hipSuccess && has15) e = hipMalloc(&pl->d_scratch, ((size_t)pl->grid << 14) * 8);
  if (e != hipSuccess) {
    plan_free(pl);
    return fail(YA_ERR_ALLOC, "batch plan allocation: %s", hipGetErrorString(e));
  }
  BatchQueue q0{};
  e = hipMemcpy(pl->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(pl->d_res, res.data(), res.size() * sizeof(BatchRes), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(pl->d_q, &q0, sizeof q0, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    plan_free(pl);
    return hip_fail(e, "batch plan upload");
  }
  pl->total_polymuls = (double)njobs;
  // Launch form: a batch of one degree class <= 2^14 with at least two
  // residues per CU runs as a grid of the fused kernel over the residue table
  // (no queue, fewer instructions per residue); mixed batches and 2^15 rings
  // use the persistent kernel.  g_form_override forces one form (tests).
  bool one_class = true;
  for (const auto &it : items) one_class &= it.logn == items[0].logn;
  const int lg = (int)items[0].logn;
  const bool grid_ok = one_class && lg <= 14;
  int form = (grid_ok && res.size() >= 2 * (size_t)cus) ? YA_BATCH_GRID : YA_BATCH_PERSISTENT;
  int k = one_class ? split_factor(st, lg, res.size()) : 0;
  if (k) form = YA_BATCH_SPLIT;
  const int ov = g_form_override.load();
  if (ov == YA_BATCH_SPLIT && one_class && !k && lg >= kSplitMinLog && lg <= kSplitMaxLog)
    k = split_choose_k(lg, res.size(), cus);
  if (ov == YA_BATCH_PERSISTENT || (ov == YA_BATCH_GRID && grid_ok) || (ov == YA_BATCH_SPLIT && k)) form = ov;
  if (form == YA_BATCH_SPLIT) {
    e = hipMalloc(&pl->d_split, (res.size() << lg) * 8);
    if (e != hipSuccess) {
      plan_free(pl);
      return fail(YA_ERR_ALLOC, "split scratch: %s", hipGetErrorString(e));
    }
    pl->split_k = k;
  }
  if (form == YA_BATCH_GRID) {
    // jobs k = 0.. at a_0 + k*N*L (and likewise b, c), one shape, c apart
    // from a and b: the direct-addressed fused kernel (no residue table to
    // read; 1.153 against 1.170 ms per 512 poly-mults at N = 2^14, L = 8 on
    // MI355X)
    const size_t w = jobs[0].degree * jobs[0].nmoduli, all = w * njobs;
    bool strided = true;
    for (size_t j = 0; j < njobs && strided; ++j)
      strided = jobs[j].degree == jobs[0].degree && jobs[j].nmoduli == jobs[0].nmoduli &&
                jobs[j].a == jobs[0].a + j * w && jobs[j].b == jobs[0].b + j * w && jobs[j].c == jobs[0].c + j * w;
    const uint64_t *c0 = jobs[0].c;
    strided = strided && !(c0 < jobs[0].a + all && jobs[0].a < c0 + all) && !(c0 < jobs[0].b + all && jobs[0].b < c0 + all);
    if (strided && g_strided.load()) {
      pl->strided = true;
      pl->sa = jobs[0].a;
      pl->sb = jobs[0].b;
      pl->sc = jobs[0].c;
      pl->s_nmod = jobs[0].nmoduli;
      pl->s_desc = desc_ptr(st, lg, 0);
    }
  }
  pl->form = form;
  pl->grid_logn = lg;
  *out = pl;
  return YA_OK;
}

static hipError_t plan_launch(ya_batch_plan *pl, hipStream_t s) {
  if (pl->form == YA_BATCH_SPLIT) {
    const SplitArgs sa{pl->d_res, nullptr, nullptr, nullptr, nullptr, 0, pl->d_split};
    return launch_polymul_split(sa, pl->grid_logn, pl->nres, pl->split_k, s);
  }
  if (pl->form == YA_BATCH_GRID && pl->strided)
    return launch_polymul(pl->sc, pl->sa, pl->sb, pl->s_desc, pl->grid_logn, pl->nres, pl->s_nmod,
                          tuned_shape(pl->grid_logn, 1), s);
  if (pl->form == YA_BATCH_GRID)
    return launch_polymul_table(pl->d_res, pl->nres, pl->grid_logn, tuned_shape(pl->grid_logn, 1), s);
  return launch_batch(pl->d_items, pl->segs, pl->d_res, pl->d_q, pl->d_scratch, pl->grid, s);
}

int ya_batch_plan_run(ya_batch_plan *pl, void *stream) {
  if (!pl) return fail(YA_ERR_INVALID, "null plan");
  hipError_t e = plan_launch(pl, (hipStream_t)stream);
  return e == hipSuccess ? YA_OK : hip_fail(e, "batch launch");
}

int ya_batch_plan_run_graph(ya_batch_plan *pl, void *stream) {
  if (!pl) return fail(YA_ERR_INVALID, "null plan");
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  if (!pl->exec) {
    hipStream_t cap;
    e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "capture stream");
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      e = plan_launch(pl, cap);
      hipError_t e2 = hipStreamEndCapture(cap, &pl->graph);
      if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cap);
    if (e != hipSuccess) return hip_fail(e, "graph capture");
  }
  e = hipGraphLaunch(pl->exec, s);
  return e == hipSuccess ? YA_OK : hip_fail(e, "graph launch");
}

int ya_batch_plan_destroy(ya_batch_plan *pl) { return plan_free(pl); }

int ya_batch_plan_info(const ya_batch_plan *pl, size_t *items, size_t *residues, size_t *workgroups,
                       int *form) {
  if (!pl) return fail(YA_ERR_INVALID, "null plan");
  if (items) *items = pl->nitems;
  if (residues) *residues = pl->nres;
  if (workgroups)
    *workgroups = pl->form == YA_BATCH_GRID    ? pl->nres
                  : pl->form == YA_BATCH_SPLIT ? pl->nres << pl->split_k
                                               : pl->grid;
  if (form) *form = pl->form;
  return YA_OK;
}

int ya_batch_plan_is_strided(const ya_batch_plan *pl, int *strided) {
  if (!pl || !strided) return fail(YA_ERR_INVALID, "null argument");
  *strided = pl->strided ? 1 : 0;
  return YA_OK;
}

int ya_batch_strided(int on) {
  g_strided.store(on ? 1 : 0);
  return YA_OK;
}

int ya_batch_force_form(int form) {
  if (form != 0 && form != YA_BATCH_PERSISTENT && form != YA_BATCH_GRID && form != YA_BATCH_SPLIT)
    return fail(YA_ERR_INVALID, "bad launch form %d", form);
  g_form_override.store(form);
  return YA_OK;
}

int ya_batch_polymul(const ya_polymul_job *jobs, size_t njobs, void *stream) {
  ya_batch_plan *pl = nullptr;
  int rc = ya_batch_plan_create(&pl, jobs, njobs);
  if (rc) return rc;
  rc = ya_batch_plan_run(pl, stream);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  plan_free(pl);
  if (rc) return rc;
  return e == hipSuccess ? YA_OK : hip_fail(e, "batch sync");
}

}  // extern "C"

// ===========================================================================
// Autotuner.  For each degree class it times every instantiated pass shape
// (LE = 3: radix-8 register passes and twice as many residues per workgroup
// tile; LE = 4: radix-16 passes) on zero-filled synthetic operands of
// 16 x 2^14 words per CU (3 x 512 MiB on 256 CUs, freed before it returns),
// and keeps the faster one in the
// launch-shape table used by later plans and fused launches.  Timing is
// data-independent (the butterflies are branch-free); each candidate's time
// is its best of kTuneRounds single launches, taken round-robin over the
// class's candidates after the clocks have settled.
namespace {

// One timed call of fn on s (ms through events a, b).
int time_once(hipStream_t s, hipEvent_t a, hipEvent_t b, float *ms, const std::function<int()> &fn) {
  (void)hipEventRecord(a, s);
  int rc = fn();
  (void)hipEventRecord(b, s);
  if (hipEventSynchronize(b) != hipSuccess) return fail(YA_ERR_LAUNCH, "tune sync");
  *ms = 0;
  (void)hipEventElapsedTime(ms, a, b);
  return rc;
}

// One candidate launch shape of a degree class and where its time goes in
// the report.
struct TuneCand {
  int slot;         // report column
  bool persistent;  // batch plan (persistent kernel) or fused grid launch
  int le, dual;
  ya_batch_plan *pl;
  float best;
};

// The GPU's clocks are low after an idle period and take ~0.1 s of work to
// ramp up; the first class's first candidate runs untimed for that long.
constexpr double kTuneSettleSeconds = 0.15;
constexpr int kTuneRounds = 5;

}  // namespace

extern "C" int ya_autotune(int logn_lo, int logn_hi, double *report) {
  DeviceState *st;
  int rc = current_device(&st);
  if (rc) return rc;
  if (logn_lo < kBatchMinLog || logn_hi > kBatchMaxLog || logn_lo > logn_hi)
    return fail(YA_ERR_INVALID, "tune range 2^%d..2^%d outside 2^%d..2^%d", logn_lo, logn_hi,
                kBatchMinLog, kBatchMaxLog);
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t L = 8;
  if ((rc = ensure_tables(st, 14, 0, L))) return rc;
  // 16 full items per CU at 2^14 (the size of a 512-poly batch of 8
  // moduli): the persistent kernel's start-up and the staggered start are
  // then weighed as in a throughput batch
  const size_t words = (size_t)cus * 16 * (1u << 14);
  uint64_t *buf = nullptr;
  hipStream_t s = nullptr;
  hipError_t e = hipMalloc(&buf, 3 * words * 8);
  if (e != hipSuccess) return fail(YA_ERR_ALLOC, "tune buffers: %s", hipGetErrorString(e));
  e = hipMemset(buf, 0, 3 * words * 8);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)hipFree(buf);
    return hip_fail(e, "tune set-up");
  }
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess)
    rc = fail(YA_ERR_LAUNCH, "tune events");
  uint64_t *a = buf, *b = buf + words, *c = buf + 2 * words;
  for (int lg = logn_lo; lg <= logn_hi && rc == YA_OK; ++lg) {
    double *rep = report ? report + YA_TUNE_REPORT_STRIDE * (lg - logn_lo) : nullptr;
    if (rep)
      for (int k = 0; k < YA_TUNE_REPORT_STRIDE; ++k) rep[k] = -1.0;
    if ((rc = ensure_tables(st, lg, 0, L))) break;
    const size_t n = (size_t)1 << lg;
    const size_t njobs = words / (n * L);
    std::vector<ya_polymul_job> jobs(njobs);
    for (size_t j = 0; j < njobs; ++j)
      jobs[j] = ya_polymul_job{a + j * n * L, b + j * n * L, c + j * n * L, (uint32_t)n, (uint32_t)L};
    const int lemin = (lg >= 14) ? 4 : 3, lemax = (lg >= 12 && lg <= 14) ? 5 : 4;
    std::vector<TuneCand> cands;
    for (int le = lemin; le <= lemax && rc == YA_OK; ++le) {
      if (le <= 4) {  // the batched kernel's 1024-thread tiles take radix 8 or 16
        set_tuned_shape(lg, 2, NttShape{le});
        ya_batch_plan *pl = nullptr;
        // time the persistent kernel itself (this uniform batch would
        // otherwise run in the grid form, which uses the fused shape)
        const int saved_form = g_form_override.exchange(YA_BATCH_PERSISTENT);
        rc = ya_batch_plan_create(&pl, jobs.data(), jobs.size());
        g_form_override.store(saved_form);
        if (rc) break;
        cands.push_back(TuneCand{le - 3, true, le, 0, pl, 1e30f});
      }
      if (lg > 14) continue;
      // the fused kernel: both forward transforms as one pass sequence with
      // shared twiddles (dual, built for radix 8 / 16 from 2^10) and as two
      // sequences, each timed where it exists
      const bool has_dual = lg >= 10 && le <= 4;
      for (int dual = has_dual ? 1 : 0; dual >= 0; --dual)
        cands.push_back(TuneCand{(has_dual && dual == 0) ? 5 + le - 3 : 2 + le - 3, false, le, dual, nullptr, 1e30f});
    }
    auto run = [&](const TuneCand &cd) -> int {
      if (cd.persistent) return ya_batch_plan_run(cd.pl, s);
      NttShape sh{cd.le};
      sh.dual = cd.dual;
      sh.stagger = tuned_stagger(lg);
      hipError_t ee = launch_polymul(c, a, b, desc_ptr(st, lg, 0), lg, njobs * L, L, sh, s);
      return ee == hipSuccess ? YA_OK : hip_fail(ee, "tune fused launch");
    };
    if (rc == YA_OK && !cands.empty() && lg == logn_lo) {
      const auto t0 = std::chrono::steady_clock::now();
      do {
        rc = run(cands[0]);
        if (rc == YA_OK && hipStreamSynchronize(s) != hipSuccess) rc = fail(YA_ERR_LAUNCH, "tune settle");
      } while (rc == YA_OK &&
               std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() < kTuneSettleSeconds);
    }
    // every candidate once untimed, then kTuneRounds rounds over all of
    // them (round-robin, so a drift of the clocks reaches each alike)
    for (auto &cd : cands)
      if (rc == YA_OK) rc = run(cd);
    for (int r = 0; r < kTuneRounds && rc == YA_OK; ++r)
      for (auto &cd : cands) {
        float ms = 0;
        if (rc == YA_OK) rc = time_once(s, ev0, ev1, &ms, [&]() { return run(cd); });
        if (ms < cd.best) cd.best = ms;
      }
    float best[2] = {1e30f, 1e30f};
    int arg[2] = {4, 4}, arg_dual = tuned_shape(lg, 1).dual;
    for (auto &cd : cands) {
      if (rep && rc == YA_OK) rep[cd.slot] = cd.best;
      const int kind = cd.persistent ? 0 : 1;
      if (cd.best < best[kind]) {
        best[kind] = cd.best, arg[kind] = cd.le;
        if (!cd.persistent) arg_dual = cd.dual;
      }
      if (cd.pl) plan_free(cd.pl);
    }
    if (rc) break;
    set_tuned_shape(lg, 2, NttShape{arg[0]});
    set_tuned_shape(lg, 1, NttShape{arg[1]});
    set_tuned_shape(lg, 0, NttShape{arg[1]});
    if (lg >= 10 && lg <= 14 && arg[1] <= 4) set_tuned_dual(lg, arg_dual);
  }
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  (void)hipStreamDestroy(s);
  (void)hipFree(buf);
  return rc;
}

extern "C" int ya_tune_get(int kind, int logn) {
  if (kind < 0 || kind > 5 || logn < 0 || logn > kMaxLogN) return -1;
  if (kind == 3) return tuned_shape(logn, 1).dual;
  if (kind == 4) return tuned_mac_block_log(logn);
  if (kind == 5) return tuned_stagger(logn);
  return tuned_shape(logn, kind).le;
}

extern "C" int ya_tune_set(int kind, int logn, int le) {
  if (kind == 3 && logn >= 0 && logn <= kMaxLogN && (le == 0 || le == 1)) {
    set_tuned_dual(logn, le);
    return YA_OK;
  }
  if (kind == 4 && logn >= 0 && logn <= kMaxLogN &&
      (le == 14 || (le == 13 && logn <= 18) || (le == 12 && logn <= 17))) {
    set_tuned_mac_block_log(logn, le);
    return YA_OK;
  }
  if (kind == 5 && logn >= 0 && logn <= kMaxLogN && le >= 0 && le <= kMaxStaggerSteps) {
    set_tuned_stagger(logn, le);
    return YA_OK;
  }
  if (kind < 0 || kind > 2 || logn < 0 || logn > kMaxLogN || le < 3 || le > 5 || (kind == 2 && le == 5))
    return fail(YA_ERR_INVALID, "bad tune entry kind=%d logn=%d le=%d", kind, logn, le);
  NttShape sh;
  sh.le = le;
  set_tuned_shape(logn, kind, sh);
  return YA_OK;
}
