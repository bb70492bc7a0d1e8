// ref_shim.cpp -- extern "C" entry points over the reference's own sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (by oracle/Makefile) together with the
// reference's lib/ntt.cpp, lib/ops.cpp, lib/math.cpp and lib/params.cpp, read
// in place under the reference checkout, into oracle/_ref/libyellref.so, and
// with -DYELL_USE_32BITS_MODULI (words T = uint32_t, degrees up to 2^15) into
// oracle/_ref/libyellref32.so.  It
// drives ntt<N>::forward / forward_lazy / backward and the ops:: functors
// directly on caller buffers (the poly<N> container is bypassed because its
// samplers are unseeded and its header pulls in GMP/MPFR).  Used to pin the
// C restatement in yell_oracle.c and as the CPU baseline of bench.py.
#include "yell/ntt.hpp"
#include "yell/ops.hpp"

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

using T = yell::params::value_type;

namespace {

// poly-mult on caller arrays, the call sequence of details/poly.hpp:171-186
// and ops.hpp:7-24: forward(a), forward(b), rop = a * b, rop.backward().
template <size_t N>
void polymul_n(T *c, const T *a, const T *b, size_t L, T *wa, T *wb) {
  yell::ops::mulmod mulmod;
  std::memcpy(wa, a, sizeof(T) * N * L);
  std::memcpy(wb, b, sizeof(T) * N * L);
  for (size_t cm = 0; cm < L; ++cm) yell::ntt<N>::forward(wa + cm * N, cm);
  for (size_t cm = 0; cm < L; ++cm) yell::ntt<N>::forward(wb + cm * N, cm);
  for (size_t cm = 0; cm < L; ++cm) {
    T *dst = c + cm * N;
    const T *x = wa + cm * N, *y = wb + cm * N;
    for (size_t i = 0; i < N; ++i) dst[i] = mulmod(x[i], y[i], cm);
  }
  for (size_t cm = 0; cm < L; ++cm) yell::ntt<N>::backward(c + cm * N, cm);
}

// The RLWE flow of the reference's example.cpp:12-58 on caller arrays, each
// [L][N], with the reference's transforms and functors in the order the
// example's poly<N> calls run them (add_product_of: details/poly.hpp:119-132,
// negate: :152-158, forward / forward_lazy / backward: :171-186, +=:
// ops.hpp:26-41).  The samples come from the caller: sx (hwt), pk_bx (the
// Gaussian e0), pk_ax (uniform), u (ZO), ct_bx (Gaussian e1), ct_ax (Gaussian
// e2); sx, pk_bx, u, ct_bx, ct_ax are transformed in place.
template <size_t N>
struct Rlwe {
  static void fwd(T *x, size_t L) { for (size_t cm = 0; cm < L; ++cm) yell::ntt<N>::forward(x + cm * N, cm); }
  static void fwd_lazy(T *x, size_t L) {
    for (size_t cm = 0; cm < L; ++cm) yell::ntt<N>::forward_lazy(x + cm * N, cm);
  }
  static void bwd(T *x, size_t L) { for (size_t cm = 0; cm < L; ++cm) yell::ntt<N>::backward(x + cm * N, cm); }
  static void add_product_of(T *acc, const T *a, const T *b, size_t L) {
    yell::ops::muladd muladd;
    for (size_t cm = 0; cm < L; ++cm)
      for (size_t i = 0; i < N; ++i) muladd.compute(acc[cm * N + i], a[cm * N + i], b[cm * N + i], cm);
  }
  static void negate(T *x, size_t L) {
    for (size_t cm = 0; cm < L; ++cm) {
      const T p = yell::params::P[cm];
      for (size_t i = 0; i < N; ++i) x[cm * N + i] = p - x[cm * N + i];
    }
  }
  static void add(T *x, const T *y, size_t L) {
    yell::ops::addmod addmod;
    for (size_t cm = 0; cm < L; ++cm)
      for (size_t i = 0; i < N; ++i) addmod.compute(x[cm * N + i], y[cm * N + i], cm);
  }
  static void keygen(size_t L, T *sx, T *pk_bx, const T *pk_ax) {
    fwd(sx, L);                         // example.cpp:15
    fwd(pk_bx, L);                      // :25
    add_product_of(pk_bx, pk_ax, sx, L);  // :26
    negate(pk_bx, L);                   // :27
  }
  static void encrypt(size_t L, const T *pk_bx, const T *pk_ax, const T *msg, T *u, T *ct_bx, T *ct_ax) {
    fwd(u, L);                          // :40
    add(ct_bx, msg, L);                 // :42
    fwd_lazy(ct_bx, L);                 // :43
    add_product_of(ct_bx, u, pk_bx, L);   // :44
    fwd_lazy(ct_ax, L);                 // :47
    add_product_of(ct_ax, u, pk_ax, L);   // :48
  }
  static void decrypt(size_t L, const T *sx, const T *ct_bx, const T *ct_ax, T *plain) {
    std::memcpy(plain, ct_bx, sizeof(T) * N * L);  // :55
    add_product_of(plain, sx, ct_ax, L);  // :56
    bwd(plain, L);                      // :57
  }
  // `reps` encrypt + decrypt rounds on each of `threads` std::threads; each
  // round first copies the samples into the thread's own buffers (about 1% of
  // a round); tables built beforehand.  Returns wall-clock seconds.
  static double bench(size_t L, const T *sx, const T *pk_bx, const T *pk_ax, const T *msg, const T *u,
                      const T *e1, const T *e2, int threads, int reps) {
    for (size_t cm = 0; cm < L; ++cm) yell::ntt<N>::init_table(cm);
    std::vector<std::thread> pool;
    auto t0 = std::chrono::steady_clock::now();
    for (int th = 0; th < threads; ++th) {
      pool.emplace_back([=]() {
        std::vector<T> wu(N * L), wb(N * L), wa(N * L), pl(N * L);
        for (int r = 0; r < reps; ++r) {
          std::memcpy(wu.data(), u, sizeof(T) * N * L);
          std::memcpy(wb.data(), e1, sizeof(T) * N * L);
          std::memcpy(wa.data(), e2, sizeof(T) * N * L);
          encrypt(L, pk_bx, pk_ax, msg, wu.data(), wb.data(), wa.data());
          decrypt(L, sx, wb.data(), wa.data(), pl.data());
        }
      });
    }
    for (auto &t : pool) t.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

template <size_t N>
struct Ops {
  static constexpr size_t kN = N;
  static void fwd(T *x, size_t cm) { yell::ntt<N>::forward(x, cm); }
  static void fwd_lazy(T *x, size_t cm) { yell::ntt<N>::forward_lazy(x, cm); }
  static void bwd(T *x, size_t cm) { yell::ntt<N>::backward(x, cm); }
  static void tables(size_t cm, T *phi, T *phis, T *iphi, T *iphis, T *sc) {
    auto t = yell::ntt<N>::init_table(cm);
    std::memcpy(phi, t->phiTbl, sizeof(T) * N);
    std::memcpy(phis, t->shoupPhiTbl, sizeof(T) * N);
    std::memcpy(iphi, t->invphiTbl, sizeof(T) * N);
    std::memcpy(iphis, t->shoupInvphiTbl, sizeof(T) * N);
    sc[0] = t->invDegree;
    sc[1] = t->shoupInvDegree;
    sc[2] = t->invphiInvDegree;
    sc[3] = t->shoupInvphiInvDegree;
  }
  static void init(size_t k) { yell::ntt<N>::init_ntt_tables(k); }
  static void polymul(T *c, const T *a, const T *b, size_t L) {
    std::vector<T> wa(N * L), wb(N * L);
    polymul_n<N>(c, a, b, L, wa.data(), wb.data());
  }
  // `reps` poly-mults on each of `threads` std::threads, each thread on its
  // own copy of the operands; read-only shared tables (built beforehand).
  // Returns the wall-clock seconds of the timed region.
  static double bench(const T *a, const T *b, size_t L, int threads, int reps) {
    init(L);
    std::vector<std::thread> pool;
    auto t0 = std::chrono::steady_clock::now();
    for (int th = 0; th < threads; ++th) {
      pool.emplace_back([=]() {
        std::vector<T> wa(N * L), wb(N * L), c(N * L);
        for (int r = 0; r < reps; ++r) polymul_n<N>(c.data(), a, b, L, wa.data(), wb.data());
      });
    }
    for (auto &t : pool) t.join();
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
  }
};

// 2^18 and 2^19 exist only in the 64-bit build (kMaxPolyDegree 2^20,
// params_64.hpp:21; the 32-bit build stops at 2^15, params_32.hpp).
#ifdef YELL_USE_32BITS_MODULI
#define YREF_LARGE_CASES(EXPR)
#else
#define YREF_LARGE_CASES(EXPR)                                                    \
    case 18: { using O = Ops<(1u << 18)>; EXPR; } break;                          \
    case 19: { using O = Ops<(1u << 19)>; EXPR; } break;
#endif

#define YREF_SWITCH(logn, EXPR)                                                   \
  switch (logn) {                                                                 \
    case 2: { using O = Ops<(1u << 2)>; EXPR; } break;                            \
    case 3: { using O = Ops<(1u << 3)>; EXPR; } break;                            \
    case 4: { using O = Ops<(1u << 4)>; EXPR; } break;                            \
    case 5: { using O = Ops<(1u << 5)>; EXPR; } break;                            \
    case 6: { using O = Ops<(1u << 6)>; EXPR; } break;                            \
    case 7: { using O = Ops<(1u << 7)>; EXPR; } break;                            \
    case 8: { using O = Ops<(1u << 8)>; EXPR; } break;                            \
    case 9: { using O = Ops<(1u << 9)>; EXPR; } break;                            \
    case 10: { using O = Ops<(1u << 10)>; EXPR; } break;                          \
    case 11: { using O = Ops<(1u << 11)>; EXPR; } break;                          \
    case 12: { using O = Ops<(1u << 12)>; EXPR; } break;                          \
    case 13: { using O = Ops<(1u << 13)>; EXPR; } break;                          \
    case 14: { using O = Ops<(1u << 14)>; EXPR; } break;                          \
    case 15: { using O = Ops<(1u << 15)>; EXPR; } break;                          \
    case 16: { using O = Ops<(1u << 16)>; EXPR; } break;                          \
    case 17: { using O = Ops<(1u << 17)>; EXPR; } break;                          \
    YREF_LARGE_CASES(EXPR)                                                        \
    default: return -1;                                                           \
  }

}  // namespace

extern "C" {

int yref_forward(int logn, T *x, size_t cm) { YREF_SWITCH(logn, O::fwd(x, cm)); return 0; }
int yref_forward_lazy(int logn, T *x, size_t cm) { YREF_SWITCH(logn, O::fwd_lazy(x, cm)); return 0; }
int yref_backward(int logn, T *x, size_t cm) { YREF_SWITCH(logn, O::bwd(x, cm)); return 0; }
int yref_tables(int logn, size_t cm, T *phi, T *phis, T *iphi, T *iphis, T *sc) {
  YREF_SWITCH(logn, O::tables(cm, phi, phis, iphi, iphis, sc));
  return 0;
}
int yref_init_tables(int logn, size_t k) { YREF_SWITCH(logn, O::init(k)); return 0; }
int yref_polymul(int logn, T *c, const T *a, const T *b, size_t L) {
  YREF_SWITCH(logn, O::polymul(c, a, b, L));
  return 0;
}
double yref_bench_polymul(int logn, const T *a, const T *b, size_t L, int threads, int reps) {
  double s = -1.0;
  switch (logn) {
    case 12: s = Ops<(1u << 12)>::bench(a, b, L, threads, reps); break;
    case 13: s = Ops<(1u << 13)>::bench(a, b, L, threads, reps); break;
    case 14: s = Ops<(1u << 14)>::bench(a, b, L, threads, reps); break;
    case 15: s = Ops<(1u << 15)>::bench(a, b, L, threads, reps); break;
    case 16: s = Ops<(1u << 16)>::bench(a, b, L, threads, reps); break;
    default: break;
  }
  return s;
}

int yref_rlwe(int logn, size_t L, T *sx, T *pk_bx, const T *pk_ax, const T *msg, T *u, T *ct_bx, T *ct_ax,
              T *plain) {
  YREF_SWITCH(logn, (Rlwe<O::kN>::keygen(L, sx, pk_bx, pk_ax), Rlwe<O::kN>::encrypt(L, pk_bx, pk_ax, msg, u, ct_bx, ct_ax),
                     Rlwe<O::kN>::decrypt(L, sx, ct_bx, ct_ax, plain)));
  return 0;
}
double yref_bench_rlwe(int logn, size_t L, const T *sx, const T *pk_bx, const T *pk_ax, const T *msg, const T *u,
                       const T *e1, const T *e2, int threads, int reps) {
  double s = -1.0;
  switch (logn) {
    case 12: s = Rlwe<(1u << 12)>::bench(L, sx, pk_bx, pk_ax, msg, u, e1, e2, threads, reps); break;
    case 13: s = Rlwe<(1u << 13)>::bench(L, sx, pk_bx, pk_ax, msg, u, e1, e2, threads, reps); break;
    case 14: s = Rlwe<(1u << 14)>::bench(L, sx, pk_bx, pk_ax, msg, u, e1, e2, threads, reps); break;
    default: break;
  }
  return s;
}

T yref_mulmod(T x, T y, size_t cm) { return yell::ops::mulmod{}(x, y, cm); }
T yref_mulmod_shoup(T x, T y, T yp, size_t cm) { return yell::ops::mulmod_shoup{}(x, y, yp, cm); }
T yref_addmod(T x, T y, size_t cm) { return yell::ops::addmod{}(x, y, cm); }
T yref_submod(T x, T y, size_t cm) { return yell::ops::submod{}(x, y, cm); }
T yref_muladd(T rop, T x, T y, size_t cm) { return yell::ops::muladd{}(rop, x, y, cm); }
T yref_shoupify(T x, size_t cm) { return yell::ops::shoupify(x, cm); }
T yref_barrett(T hi, T lo, size_t cm) {
  yell::params::gt_value_type v = ((yell::params::gt_value_type)hi << (8 * sizeof(T))) | lo;
  yell::ops::barret_reduction(&v, cm);
  return (T)v;
}

}  // extern "C"
