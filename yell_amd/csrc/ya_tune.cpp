// ya_tune.cpp -- the launch-shape table consulted by every NTT launch.  The
// autotuner (ya_autotune in ya_batch.hip) times the candidate shapes of a
// degree class on the device and stores the winner here.
#include <atomic>
#include <mutex>

#include "ya_internal.hpp"

namespace ya {

static std::atomic<int> g_le[3][kMaxLogN + 1];
static std::atomic<int> g_dual[kMaxLogN + 1];  // fused kernel: dual-forward form
static std::atomic<int> g_mac_lb[kMaxLogN + 1];  // large-ring MAC: pair-product block log
static std::atomic<int> g_stagger[kMaxLogN + 1];  // first-wave stagger steps
static std::once_flag g_init;

// Defaults, set once; a thread that arrives while another sets them waits.
static void init_defaults() {
  std::call_once(g_init, [] {
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l <= kMaxLogN; ++l) g_le[k][l].store(4);
    // dual-forward fused form: measured faster at every class it is built for
    for (int l = 0; l <= kMaxLogN; ++l) g_dual[l].store(l >= 10 ? 1 : 0);
    // large-ring MAC pair pass: 2^12 blocks where the top pass takes the
    // rest (four 256-thread workgroups per CU; 2^16 x 24 primes, 8 pairs:
    // 0.206 ms against 0.230 with 2^13 and 0.266 with 2^14 blocks), 2^13
    // at 2^18, 2^14 at 2^19 (at most 5 top layers); up to 2^14 the ring is
    // one block
    for (int l = 0; l <= kMaxLogN; ++l) g_mac_lb[l].store(l <= 14 ? 14 : l <= 17 ? 12 : l == 18 ? 13 : 14);
    // measured faster where one workgroup fills a CU: the fused kernels at
    // 2^14 (one workgroup per residue: best at 1..3 steps; the looping
    // kernel: 0..2 at 512 polys, 2..3 at 1,024) and the persistent kernel
    // (its 2^15 entry; best at 6 steps on the uneven 1024-job batch)
    for (int l = 0; l <= kMaxLogN; ++l) g_stagger[l].store(l == 14 ? 2 : l == 15 ? 6 : 0);
  });
}

int tuned_mac_block_log(int logn) {
  init_defaults();
  const int lb = logn >= 0 && logn <= kMaxLogN ? g_mac_lb[logn].load() : 14;
  return logn <= lb ? logn : lb;
}

void set_tuned_mac_block_log(int logn, int lb) {
  init_defaults();
  if (logn >= 0 && logn <= kMaxLogN) g_mac_lb[logn].store(lb);
}

int tuned_stagger(int logn) {
  init_defaults();
  return logn >= 0 && logn <= kMaxLogN ? g_stagger[logn].load() : 0;
}

void set_tuned_stagger(int logn, int steps) {
  init_defaults();
  if (logn >= 0 && logn <= kMaxLogN) g_stagger[logn].store(steps);
}

NttShape tuned_shape(int logn, int kind) {
  init_defaults();
  NttShape s;
  if (logn >= 0 && logn <= kMaxLogN && kind >= 0 && kind < 3) s.le = g_le[kind][logn].load();
  if (kind == 1 && logn >= 0 && logn <= kMaxLogN) {
    s.dual = g_dual[logn].load();
    s.stagger = g_stagger[logn].load();
  }
  return s;
}

void set_tuned_shape(int logn, int kind, NttShape s) {
  init_defaults();
  if (logn >= 0 && logn <= kMaxLogN && kind >= 0 && kind < 3) g_le[kind][logn].store(s.le);
}

void set_tuned_dual(int logn, int dual) {
  init_defaults();
  if (logn >= 0 && logn <= kMaxLogN) g_dual[logn].store(dual ? 1 : 0);
}

}  // namespace ya
