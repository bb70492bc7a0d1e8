// ya_internal.hpp -- declarations shared by the host-side translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "ya_kernels.hpp"

namespace ya {

struct DeviceState;

// Records a formatted message for ya_last_error() and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int hip_fail(hipError_t e, const char *what);

// State of the calling thread's current HIP device (created on first use).
int current_device(DeviceState **out);
// Makes the tables of (2^logn, cm) for cm in [cm0, cm0+count) device-resident.
int ensure_tables(DeviceState *st, int logn, size_t cm0, size_t count);
// Device descriptor of (2^logn, cm); descriptors of consecutive cm are adjacent.
const ModDesc *desc_ptr(DeviceState *st, int logn, size_t cm);
// Base of the device descriptor array: entry (logn, cm) is at logn * 1000 + cm.
const ModDesc *desc_table_base(DeviceState *st);

// Compute units of the device (256 on MI355X).
int device_cus(DeviceState *st);
// Split factor for a poly-multiplication launch of nres residues of 2^logn
// words under the split policy (ya_split_policy); 0 = one workgroup per residue.
int split_factor(DeviceState *st, int logn, size_t nres);

int log2_degree(size_t degree);
int check_shape(size_t degree, size_t cm0, size_t nmoduli, int *logn);
uint64_t prime(size_t cm);

// Tuned launch shape for a degree class.  kind 0: standalone NTT kernels,
// kind 1: fused poly-multiplication, kind 2: batched poly-multiplication.
NttShape tuned_shape(int logn, int kind);
void set_tuned_shape(int logn, int kind, NttShape s);  // sets le only
// Fused poly-multiplication of degree 2^logn: dual-forward form on/off.
void set_tuned_dual(int logn, int dual);
// Large-ring MAC (ya_poly_add_products_backward) above 2^13: block log of
// its pair-product pass, 14 (one 1024-thread workgroup per CU), 13 (two
// 512-thread workgroups) or 12 (four 256-thread workgroups); the result
// never depends on it.
int tuned_mac_block_log(int logn);
void set_tuned_mac_block_log(int logn, int lb);
// Staggered start of the first wave of workgroups (stagger_start in
// ya_blocks.hpp): steps of ~0.85 us per group for the fused kernel of degree
// 2^logn, and for the persistent kernel at logn = 15; 0 = off.
constexpr int kMaxStaggerSteps = 16;
int tuned_stagger(int logn);
void set_tuned_stagger(int logn, int steps);

// Pinned host memory registered for kernel access (ya_host_alloc_pinned
// allocates through pinned_alloc).  pinned_device_range returns the device
// address of [p, p + bytes) when the whole range lies inside one registered
// allocation, else null.
int pinned_alloc(void **ptr, size_t bytes);
int pinned_free(void *ptr);
void *pinned_device_range(const void *p, size_t bytes);

// Largest Gaussian parameter of the samplers: ceil(13 sigma) stays below every
// prime of both builds, so a signed sample maps to one residue without reduction.
constexpr double kMaxGaussianSigma = 1e6;
// Cumulative table of the Gaussian samplers (host and device use the same one).
int gaussian_cdt(double sigma, std::vector<uint64_t> &cdf);

}  // namespace ya
