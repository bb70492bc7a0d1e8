// ya_tune.cpp -- the launch-shape table consulted by every NTT launch.  The
// autotuner (ya_autotune in ya_batch.hip) times the candidate shapes of a
// degree class on the device and stores the winner here.
#include <atomic>

#include "ya_internal.hpp"

namespace ya {

static std::atomic<int> g_le[3][kMaxLogN + 1];
static std::atomic<int> g_dual[kMaxLogN + 1];  // fused kernel: dual-forward form
static std::atomic<bool> g_init{false};

static void init_defaults() {
  bool expected = false;
  if (!g_init.compare_exchange_strong(expected, true)) return;
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l <= kMaxLogN; ++l) g_le[k][l].store(4);
  // dual-forward fused form: measured faster at every class it is built for
  for (int l = 0; l <= kMaxLogN; ++l) g_dual[l].store(l >= 10 ? 1 : 0);
}

NttShape tuned_shape(int logn, int kind) {
  init_defaults();
  NttShape s;
  if (logn >= 0 && logn <= kMaxLogN && kind >= 0 && kind < 3) s.le = g_le[kind][logn].load();
  if (kind == 1 && logn >= 0 && logn <= kMaxLogN) s.dual = g_dual[logn].load();
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
