// ya_kernels.hpp -- host-side launch wrappers of the HIP kernels (C++ only,
// no templates leak out).  All return hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ya_ntt.hpp"

namespace ya {

// Largest block handled in one workgroup's LDS (2^14 words = 128 KiB + padding).
constexpr int kMaxBlockLog = 14;
// Degrees 2^2 .. 2^19: the reference's poly<N> requires N < kMaxPolyDegree =
// 2^20 (include/yell/poly.hpp:45, params_64.hpp:21), and its forward loop
// reads past a residue of 2 words (lib/ntt.cpp:123-135 assumes h = 2 exists).
constexpr int kMinLogN = 2;
constexpr int kMaxLogN = 19;

// Elements per thread (log2) chosen for a block of 2^lb words; 0 = default.
struct NttShape {
  int le = 4;
  int dual = 0;     // fused poly-multiplication: both forward transforms in one pass sequence
  int stagger = 0;  // fused poly-multiplication: first-wave stagger steps (stagger_start)
};

// forward / forward_lazy / backward over `nres` residues of 2^logn words,
// residue r at data + r * 2^logn with descriptor descs[r % nmod] (nmod = nres
// for one polynomial; nmod = L for nres / L polynomials of L residues).
hipError_t launch_forward(uint64_t *data, const ModDesc *descs, int logn, size_t nres, size_t nmod,
                          bool lazy, NttShape shape, hipStream_t st);
// Forward transforms of polynomials 0 .. 2*npairs-1 of a many-polynomial
// buffer (nmod residues each), two same-prime residues per workgroup; degrees
// 2^kMinPairLog .. 2^kMaxLogN with radix 8 / 16 passes (shape.le <= 4).
// Above 2^kMaxBlockLog the top layers run per residue first (k_fwd_top), then
// block b of residue r of polynomials 2q and 2q+1 share a workgroup.
// Measured on MI355X: 3-6% faster than one residue per workgroup at 2^13 and
// 2^14 (LDS-bound there), 9-19% slower at 2^10..2^12 (the single kernel runs
// more waves per SIMD in fewer registers).
constexpr int kMinPairLog = 13;
hipError_t launch_forward_pairs(uint64_t *data, const ModDesc *descs, int logn, size_t npairs, size_t nmod,
                                bool lazy, NttShape shape, hipStream_t st);
bool forward_pairs_ok(int logn, NttShape shape);
hipError_t launch_backward(uint64_t *data, const ModDesc *descs, int logn, size_t nres, size_t nmod,
                           NttShape shape, hipStream_t st);

// Fused steps of the reference example's RLWE flow, the words of the unfused
// launches.  acc_k = u_k * key + forward_lazy(acc_k + add_k) for polynomials
// 0 .. 2*npairs-1 (add may be null; key one polynomial shared by all), two
// same-prime residues per workgroup, degrees 2^kMinPairLog .. 2^kMaxBlockLog
// with shape.le <= 4; add, u and key must not alias acc.
hipError_t launch_forward_pairs_mac(uint64_t *acc, const uint64_t *add, const uint64_t *u, const uint64_t *key,
                                    const ModDesc *descs, int logn, size_t npairs, size_t nmod, NttShape shape,
                                    hipStream_t st);
// out_r = backward(acc_r + x_r * key_(r mod nmod)) for nres residues, degrees
// up to 2^kMaxBlockLog; out may be acc.
hipError_t launch_mac_inverse(uint64_t *out, const uint64_t *acc, const uint64_t *x, const uint64_t *key,
                              const ModDesc *descs, int logn, size_t nres, size_t nmod, NttShape shape,
                              hipStream_t st);

// negacylic_backward_lazy (lib/ntt.cpp:151-210): every inverse layer except
// the top one, outputs in [0, 2p).
hipError_t launch_inverse_lazy(uint64_t *data, const ModDesc *descs, int logn, size_t nres,
                               hipStream_t st);

// c = backward(forward(a) (*) forward(b)) for `nres` residues (a, b, c share
// the residue-major layout); residue r uses descs[r % nmod] (nres is a
// multiple of nmod: nres / nmod polynomials of nmod residues each); c may not
// alias a or b.  Rings up to 2^14 only.
hipError_t launch_polymul(uint64_t *c, const uint64_t *a, const uint64_t *b, const ModDesc *descs,
                          int logn, size_t nres, size_t nmod, NttShape shape, hipStream_t st);

// acc = acc + sum_j a[j] * b[j] (mod p), the pairs applied in order j = 0..n-1
// with one Barrett muladd each (details/poly.hpp:119-132 per pair); at most
// kMacPairs pairs per launch.
constexpr int kMacPairs = 16;
struct MacPairs {
  const uint64_t *a[kMacPairs];
  const uint64_t *b[kMacPairs];
};
hipError_t launch_mac(uint64_t *acc, const MacPairs &pairs, int npairs, const ModDesc *descs,
                      size_t n, size_t nres, hipStream_t st);

// The large-ring multiply-accumulate: acc.add_product_of(forward(a_j),
// forward(b_j)) for j < npairs, then acc.backward() (details/poly.hpp:119-132,
// 171-186), the words of that sequence.  Two launches groups:
// launch_pair_products: forward(a_j), forward(b_j) and their product, one
// workgroup per (pair, residue, block) with both operands' passes sharing
// twiddles; the product (canonical, natural order) goes to a_j, and b_j is
// work space (above 2^kMaxBlockLog the top layers of a_j and b_j run first,
// in place).  launch_sum_backward: acc + sum_j a_j (addmod: the same residues
// as the muladd chain on canonical words), then the inverse; with accumulate
// false acc is only written (sum_j a_j, npairs >= 1).  Degrees
// 2^kMinPairLog .. 2^kMaxLogN; at most kMacPairs pairs per launch; the 2 *
// npairs + 1 buffers must not overlap.
struct PairPtrs {
  uint64_t *a[kMacPairs];
  uint64_t *b[kMacPairs];
};
hipError_t launch_pair_products(const PairPtrs &pp, int npairs, const ModDesc *descs, int logn, size_t nmod,
                                int lb, hipStream_t st);
hipError_t launch_sum_backward(uint64_t *acc, const PairPtrs &pp, int npairs, const ModDesc *descs, int logn,
                               size_t nmod, bool accumulate, hipStream_t st);

// Copies between pinned host memory and device memory by a kernel (the
// served form's operand upload and result download: one launch for a whole
// batch instead of one DMA copy per buffer).  Each descriptor moves `words`
// 64-bit words, an even number, from src to dst, both 16-byte aligned device
// addresses (a pinned host buffer's device address from pinned_device_range).
// The descriptor array itself must be readable by the device (pinned host or
// device memory).  At most 65535 descriptors; max_words the largest `words`.
struct XferDesc {
  const uint64_t *src;
  uint64_t *dst;
  uint64_t words;
};
hipError_t launch_xfer(const XferDesc *descs, size_t n, size_t max_words, hipStream_t st);

enum PwOp { PW_MUL = 0, PW_ADD = 1, PW_SUB = 2, PW_NEGATE = 3, PW_MULADD = 4, PW_CORRECT = 5 };

// Element-wise op over `nres` residues of `n` words; residue r uses descs[r].p / pn.
//   MUL/ADD/SUB:  out = a op b;  NEGATE: out = p - a;  MULADD: out = out + a*b;
//   CORRECT: out = a reduced from [0,4p) to [0,p).
hipError_t launch_pointwise(PwOp op, uint64_t *out, const uint64_t *a, const uint64_t *b,
                            const ModDesc *descs, size_t n, size_t nres, hipStream_t st);
// `count` polynomials of nmod residues back to back in out and a; b's
// polynomial k starts at k * b_stride (0: one b for all); nmod * count <= 65535.
hipError_t launch_pointwise_many(PwOp op, uint64_t *out, const uint64_t *a, const uint64_t *b,
                                 const ModDesc *descs, size_t n, size_t nmod, size_t count, size_t b_stride,
                                 hipStream_t st);

}  // namespace ya
