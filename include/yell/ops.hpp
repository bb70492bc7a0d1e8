// yell/ops.hpp -- drop-in for the reference's scalar modular operations
// (include/yell/ops.hpp:43-158, lib/ops.cpp:5-27): the same functor names,
// signatures and results, evaluated on the host.  The element-wise poly
// operators of the reference's ops.hpp:7-41 live in yell/poly.hpp and run on
// the device.
#pragma once
#include <cstddef>
#include <type_traits>

#include "yell/params.hpp"

namespace yell {
namespace ops {

template <typename T>
static inline void mod_correct(T &a, const T p) {
  a -= (a >= p) ? p : T(0);
}

// Barrett (lib/ops.cpp:5-21): with w = kModulusRepresentationBitsize and
// d = w - kModulusBitsize, q = Pn * (x >> w) + (x << d) (mod 2^2w);
// r = lo(x) - hi(q) * p (mod 2^w); one conditional subtraction.
inline void barret_reduction(params::gt_value_type *x, size_t cm) {
  if (!x) return;
  using gT = params::gt_value_type;
  using T = params::value_type;
  constexpr unsigned w = params::kModulusRepresentationBitsize;
  constexpr unsigned d = w - params::kModulusBitsize;
  const gT v = *x;
  const gT q = (gT)params::Pn[cm] * (v >> w) + (v << d);
  T r = (T)v - (T)(q >> w) * params::P[cm];
  mod_correct(r, params::P[cm]);
  *x = r;
}

inline params::value_type shoupify(params::value_type x, size_t cm) {
  return (params::value_type)(((params::gt_value_type)x << params::kModulusRepresentationBitsize) /
                              params::P[cm]);
}

struct addmod {
  using T = params::value_type;
  T operator()(T x, T y, size_t cm) const { compute(x, y, cm); return x; }
  void compute(T &x, T y, size_t cm) const { x += y; mod_correct(x, params::P[cm]); }
};

struct submod {
  using T = params::value_type;
  T operator()(T x, T y, size_t cm) const { compute(x, y, cm); return x; }
  void compute(T &x, T y, size_t cm) const { x += params::P[cm] - y; mod_correct(x, params::P[cm]); }
};

struct mulmod {
  using T = params::value_type;
  T operator()(T x, T y, size_t cm) const { compute(x, y, cm); return x; }
  void compute(T &x, T y, size_t cm) const {
    params::gt_value_type r = (params::gt_value_type)x * y;
    barret_reduction(&r, cm);
    x = (T)r;
  }
};

struct mulmod_shoup {
  using T = params::value_type;
  T operator()(T x, T y, T yprime, size_t cm) const { compute(x, y, yprime, cm); return x; }
  void compute(T &x, T y, T yprime, size_t cm) const {
    const T p = params::P[cm];
    const T q = (T)(((params::gt_value_type)x * yprime) >> params::kModulusRepresentationBitsize) * p;
    x = x * y - q;
    mod_correct(x, p);
  }
};

struct muladd {
  using T = params::value_type;
  T operator()(T rop, T x, T y, size_t cm) const { compute(rop, x, y, cm); return rop; }
  void compute(T &rop, T x, T y, size_t cm) const {
    params::gt_value_type r = (params::gt_value_type)x * y + rop;
    barret_reduction(&r, cm);
    rop = (T)r;
  }
};

}  // namespace ops
}  // namespace yell
