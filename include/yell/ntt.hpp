// yell/ntt.hpp -- drop-in for the reference's yell::ntt<N>
// (include/yell/ntt.hpp:6-66): the same static interface, with the tables
// resident on the GPU and the transforms run by libyell_amd (yell_amd.h);
// 64-bit moduli by default, 32-bit with -DYELL_USE_32BITS_MODULI.
// op may be device memory (a poly's ptr_at) -- transformed in place on the
// device -- or host memory, which is staged through a device buffer.
// Failures throw std::runtime_error (the reference asserts).
#pragma once
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "yell/params.hpp"
#include "yell_amd.h"

namespace yell {

inline void ya_check(int rc, const char *what) {
  if (rc != YA_OK)
    throw std::runtime_error(std::string(what) + ": " + ya_status_string(rc) + ": " + ya_last_error());
}

inline bool is_device_ptr(const void *p) {
  int d = 0;
  ya_check(ya_pointer_is_device(p, &d), "ya_pointer_is_device");
  return d != 0;
}

// Entry points of the build that params::value_type selects: uint64_t words
// go to ya_*, uint32_t words (YELL_USE_32BITS_MODULI) to ya32_*.
namespace detail {
inline int api_forward(uint64_t *x, size_t n, size_t cm) { return ya_ntt_forward(x, n, cm, nullptr); }
inline int api_forward(uint32_t *x, size_t n, size_t cm) { return ya32_ntt_forward(x, n, cm, nullptr); }
inline int api_forward_lazy(uint64_t *x, size_t n, size_t cm) { return ya_ntt_forward_lazy(x, n, cm, nullptr); }
inline int api_forward_lazy(uint32_t *x, size_t n, size_t cm) { return ya32_ntt_forward_lazy(x, n, cm, nullptr); }
inline int api_backward(uint64_t *x, size_t n, size_t cm) { return ya_ntt_backward(x, n, cm, nullptr); }
inline int api_backward(uint32_t *x, size_t n, size_t cm) { return ya32_ntt_backward(x, n, cm, nullptr); }
inline int api_init_tables(const uint64_t *, size_t n, size_t k) { return ya_ntt_init_tables(n, k); }
inline int api_init_tables(const uint32_t *, size_t n, size_t k) { return ya32_ntt_init_tables(n, k); }
inline int api_table(size_t n, size_t cm, uint64_t *a, uint64_t *b, uint64_t *c, uint64_t *d, uint64_t *sc) {
  return ya_ntt_table_download(n, cm, a, b, c, d, sc);
}
// the 32-bit build's device tables are uploaded from this same host construction
inline int api_table(size_t n, size_t cm, uint32_t *a, uint32_t *b, uint32_t *c, uint32_t *d, uint32_t *sc) {
  return ya32_ntt_table_host(n, cm, a, b, c, d, sc);
}
}  // namespace detail

template <size_t degree_>
class ntt {
  static constexpr size_t degree = degree_;
  using T = params::value_type;

 public:
  // The reference's ntt_precomputed (include/yell/ntt.hpp:23-45, filled by
  // details/ntt.hpp:27-77): a host copy of the device-resident table of
  // (degree, cm), the same members as fixed arrays in bit-reversed order.
  struct ntt_precomputed {
    void init(size_t cm) {
      T sc[4];
      ya_check(detail::api_table(degree, cm, phiTbl, shoupPhiTbl, invphiTbl, shoupInvphiTbl, sc), "ntt table");
      invDegree = sc[0];
      shoupInvDegree = sc[1];
      invphiInvDegree = sc[2];
      shoupInvphiInvDegree = sc[3];
    }
    T phiTbl[degree], shoupPhiTbl[degree], invphiTbl[degree], shoupInvphiTbl[degree];
    T invDegree, shoupInvDegree, invphiInvDegree, shoupInvphiInvDegree;
  };

  static const ntt_precomputed *init_table(size_t cm) {
    static std::mutex mu;
    static std::map<size_t, std::unique_ptr<ntt_precomputed>> cache;
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(cm);
    if (it != cache.end()) return it->second.get();
    std::unique_ptr<ntt_precomputed> t(new ntt_precomputed());
    t->init(cm);
    auto *raw = t.get();
    cache.emplace(cm, std::move(t));
    return raw;
  }

  static void init_ntt_tables(size_t K) {
    ya_check(detail::api_init_tables((const T *)nullptr, degree, K), "init_ntt_tables");
  }

  static void forward(T *op, size_t cm) { run(op, cm, 0); }
  static void forward_lazy(T *op, size_t cm) { run(op, cm, 1); }
  static void backward(T *op, size_t cm) { run(op, cm, 2); }

 private:
  static void run(T *op, size_t cm, int dir) {
    auto call = [&](T *x) {
      if (dir == 0) ya_check(detail::api_forward(x, degree, cm), "ntt forward");
      else if (dir == 1) ya_check(detail::api_forward_lazy(x, degree, cm), "ntt forward_lazy");
      else ya_check(detail::api_backward(x, degree, cm), "ntt backward");
    };
    if (is_device_ptr(op)) {
      call(op);
      return;
    }
    void *d = nullptr;
    ya_check(ya_malloc(&d, degree * sizeof(T)), "ya_malloc");
    std::unique_ptr<void, int (*)(void *)> guard(d, ya_free);
    ya_check(ya_memcpy_htod(d, op, degree * sizeof(T), nullptr), "ya_memcpy_htod");
    call(static_cast<T *>(d));
    ya_check(ya_memcpy_dtoh(op, d, degree * sizeof(T), nullptr), "ya_memcpy_dtoh");
  }
};

}  // namespace yell
