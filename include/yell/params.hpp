// yell/params.hpp -- drop-in for the reference's include/yell/params.hpp:
// params_64.hpp by default, params_32.hpp with -DYELL_USE_32BITS_MODULI (the
// reference's compile-time switch, CMakeLists.txt:19).  The same names, types
// and constant tables (generated by yell_amd/csrc/gen_params.py; identical to
// the reference's values).
#pragma once
#include <cstddef>
#include <cstdint>

namespace yell {

#ifdef YELL_USE_32BITS_MODULI
struct params {
  using uint128_t = unsigned __int128;
  using value_type = uint32_t;
  using signed_type = int32_t;
  using gt_value_type = uint64_t;
  using float_type = double;

  static constexpr unsigned int kMaxNbModuli = 291;
  static constexpr unsigned int kMaxPolyDegree = 32768;
  static constexpr unsigned int kModulusBitsize = 30;
  static constexpr unsigned int kModulusRepresentationBitsize = 32;

#include "../../yell_amd/csrc/params32_table.inc"
  static constexpr value_type P[kMaxNbModuli] = YA32_TABLE_P;
  static constexpr value_type Pn[kMaxNbModuli] = YA32_TABLE_PN;
  static constexpr value_type primitive_roots[kMaxNbModuli] = YA32_TABLE_ROOT;
  static constexpr value_type invkMaxPolyDegree[kMaxNbModuli] = YA32_TABLE_INVDEG;
#undef YA32_TABLE_P
#undef YA32_TABLE_PN
#undef YA32_TABLE_ROOT
#undef YA32_TABLE_INVDEG
};
#else
struct params {
  using uint128_t = unsigned __int128;
  using value_type = uint64_t;
  using signed_type = int64_t;
  using gt_value_type = uint128_t;
  using float_type = double;

  static constexpr unsigned int kMaxNbModuli = 1000;
  static constexpr unsigned int kMaxPolyDegree = 1048576;
  static constexpr unsigned int kModulusBitsize = 62;
  static constexpr unsigned int kModulusRepresentationBitsize = 64;
  static constexpr unsigned int kMaxLazyReduction = 32;

#include "../../yell_amd/csrc/params_table.inc"
  static constexpr value_type P[kMaxNbModuli] = YA_TABLE_P;
  static constexpr value_type Pn[kMaxNbModuli] = YA_TABLE_PN;
  static constexpr value_type primitive_roots[kMaxNbModuli] = YA_TABLE_ROOT;
  static constexpr value_type invkMaxPolyDegree[kMaxNbModuli] = YA_TABLE_INVDEG;
#undef YA_TABLE_P
#undef YA_TABLE_PN
#undef YA_TABLE_ROOT
#undef YA_TABLE_INVDEG
};
#endif

}  // namespace yell
