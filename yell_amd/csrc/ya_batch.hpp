// ya_batch.hpp -- work-item layout of the batched poly-multiplication kernel
// and the host helpers shared by batch plans and the request server.
#pragma once

#include <vector>

#include "../../include/yell_amd.h"
#include "ya_internal.hpp"

namespace ya {

constexpr int kBatchThreads = 1024;
constexpr int kBatchMinLog = 7;
constexpr int kBatchMaxLog = 15;

struct BatchItem {
  uint32_t logn;   // degree class
  uint32_t le;     // log2 words per thread in a pass (3 or 4)
  uint32_t first;  // first residue (index into the residue table)
  uint32_t count;  // residues in this item
};

struct BatchRes {
  const uint64_t *a;
  const uint64_t *b;
  uint64_t *c;
  const ModDesc *d;
};

// Work items are split into kSegs segments, one per XCD (8 on MI355X):
// workgroup w drains segment w % kSegs first -- the hardware places
// consecutive workgroups on consecutive XCDs, and each segment holds whole
// (degree, prime) classes, so an XCD's L2 keeps few twiddle tables -- then
// takes items from the other segments.
constexpr int kSegs = 8;

struct BatchQueue {
  uint32_t next[kSegs];  // next item to claim in each segment (relative)
  uint32_t done;         // workgroups that have drained every segment
};

struct BatchSegs {
  uint32_t begin[kSegs + 1];  // segment s = items [begin[s], begin[s+1])
};

// Groups the residues of `jobs` into work items (sub-grouped by the tuned
// pass shape), deals whole (degree, prime) classes to the XCD segments by
// estimated work (largest first, to the least loaded segment) and orders
// each segment largest degree first; validates every job.
int build_batch(DeviceState *st, const ya_polymul_job *jobs, size_t njobs,
                std::vector<BatchItem> &items, std::vector<BatchRes> &res, BatchSegs *segs,
                bool *has15);

// Launches the persistent kernel over device copies of items / res.  `q`
// must be zeroed before the first launch (the kernel leaves it zeroed);
// `scratch` needs grid * 2^15 words when any item is of degree 2^15.
hipError_t launch_batch(const BatchItem *items, const BatchSegs &segs, const BatchRes *res,
                        BatchQueue *q, uint64_t *scratch, unsigned grid, hipStream_t s);

// Grid launch of the fused poly-multiplication over a residue table whose
// residues are all of degree 2^logn (<= 2^14), one workgroup per residue.
// Poly-multiplication of few residues, each split over 2^k workgroups
// (ya_kernels_split.hip), for degrees 2^kSplitMinLog .. 2^kSplitMaxLog.
// Residues come from a table (tab != nullptr, one degree class) or are
// contiguous (residue v at word v << logn of a, b and c, using
// descs[v % nmod]).  The top layers of forward(b) go to scratch t (residue v
// at word v << logn) or, when t is null, in place into b.
struct SplitArgs {
  const BatchRes *tab;
  const uint64_t *a;
  const uint64_t *b;
  uint64_t *c;
  const ModDesc *descs;
  uint32_t nmod;
  uint64_t *t;
};
constexpr int kSplitMinLog = 11, kSplitMaxLog = 14;
constexpr int kSplitMinK = 2, kSplitMaxK = 5, kSplitMinS = 9;
// Split factor for nres residues of 2^logn words on `cus` CUs: the smallest
// k giving at least one block per CU (within the instantiated shapes); 0 if
// the degree has no split shape.
int split_choose_k(int logn, size_t nres, int cus);
hipError_t launch_polymul_split(const SplitArgs &s, int logn, size_t nres, int k, hipStream_t st);
// A ring above 2^14 multiplies as a split launch too: blocks of 2^13 (two
// 512-thread workgroups per CU) up to 2^18, 2^14 at 2^19; returns its k.
int large_ring_split_k(int logn);

hipError_t launch_polymul_table(const BatchRes *res, size_t nres, int logn, NttShape shape,
                                hipStream_t st);

}  // namespace ya
