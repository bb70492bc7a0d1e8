// ya32.hpp -- launch wrappers of the 32-bit moduli build (ya32_kernels.hip).
#pragma once

#include "ya_kernels.hpp"

namespace ya {

constexpr int kMinLogN32 = 3;
constexpr int kMaxLogN32 = 15;  // params_32.hpp kMaxPolyDegree = 2^15

hipError_t launch32_forward(uint32_t *data, const ModDesc32 *descs, int logn, size_t nres, size_t nmod,
                            bool lazy, hipStream_t st);
// inverse with the merged N^-1 layer (canonical outputs)
hipError_t launch32_backward(uint32_t *data, const ModDesc32 *descs, int logn, size_t nres, size_t nmod,
                             hipStream_t st);
hipError_t launch32_polymul(uint32_t *c, const uint32_t *a, const uint32_t *b, const ModDesc32 *descs,
                            int logn, size_t nres, size_t nmod, hipStream_t st);
// words of residue r use descs[r % nmod]; PW_MULADD accumulates into out
hipError_t launch32_pointwise(PwOp op, uint32_t *out, const uint32_t *a, const uint32_t *b,
                              const ModDesc32 *descs, int logn, size_t nmod, size_t words, hipStream_t st);

}  // namespace ya
