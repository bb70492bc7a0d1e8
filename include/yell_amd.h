/*
 * yell_amd.h -- C ABI of the MI355X (gfx950) negacyclic NTT / RNS
 * polynomial-arithmetic engine.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns
 * YA_OK (0) or a positive YA_ERR_* code; ya_last_error() gives the detail of
 * the calling thread's last failure.  No exception crosses this boundary.
 *
 * Layout.  A polynomial with L residues of degree N is residue-major: residue
 * r occupies words [r*N, (r+1)*N) of one device buffer and is reduced modulo
 * the prime P[r] (the reference's poly<N> keeps residue r in its own buffer,
 * include/yell/poly.hpp:46-47, and uses P[r] for it, poly.hpp:117).  Device
 * buffers come from ya_malloc (or any hipMalloc'd memory).  `stream` is a
 * hipStream_t passed as void*; NULL is the default stream.  Calls taking a
 * stream are asynchronous with respect to the host.
 *
 * Value ranges follow the reference exactly: forward / backward / mul / add /
 * sub / add_product_of return canonical residues in [0, p); forward_lazy
 * returns the reference's exact representative in [0, 4p); negate maps 0 to p.
 * Preconditions (as the reference, lib/ntt.cpp:45-46,66-68, ops.hpp:64,79):
 * forward inputs < 4p, backward inputs < 2p, add/sub inputs < p.
 *
 * Supported sizes: degree N = 2^k with 2 <= k <= 19 (the reference's poly<N>
 * takes N < kMaxPolyDegree = 2^20, include/yell/poly.hpp:45,
 * params_64.hpp:21; its forward loop needs N >= 4, lib/ntt.cpp:123-135;
 * at N = 4 ntt<N>::backward returns the reference's own words, which apply
 * the top layer twice, lib/ntt.cpp:161-192);
 * nmoduli L >= 1 with (first prime index + L) <= ya_num_moduli() (1000).
 */
#ifndef YELL_AMD_H
#define YELL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  YA_OK = 0,
  YA_ERR_INVALID = 1,      /* bad size, degree, modulus index or pointer */
  YA_ERR_NO_DEVICE = 2,    /* no HIP device / HIP runtime failure at init */
  YA_ERR_ALLOC = 3,        /* device or host allocation failed */
  YA_ERR_LAUNCH = 4,       /* kernel launch or HIP call failed */
  YA_ERR_UNSUPPORTED = 5,  /* valid request this build does not handle */
  YA_ERR_REJECTED = 6,     /* server admission refused (too large / closed) */
  YA_ERR_TIMEOUT = 7,      /* wait timed out */
  YA_ERR_MISMATCH = 8      /* caller tables do not match the resident ones */
};

/* ---- status / device --------------------------------------------------- */
const char *ya_status_string(int status);
const char *ya_last_error(void);
/* Selects the HIP device for this thread's later calls and creates its state. */
int ya_init(int device);
int ya_device_count(int *count);
int ya_device_synchronize(void);

/* ---- constants (include/yell/params_64.hpp:19-34) ----------------------- */
/* params::kMaxNbModuli */
size_t ya_num_moduli(void);
/* params::P[cm], params::Pn[cm], params::primitive_roots[cm]; 0 if cm is out of range. */
uint64_t ya_prime(size_t cm);
uint64_t ya_barrett_const(size_t cm);
uint64_t ya_root(size_t cm);

/* ---- NTT tables (include/yell/ntt.hpp:50-56, details/ntt.hpp:27-104) ---- */
/* ntt<N>::init_ntt_tables(K): builds and makes device-resident the tables of
 * (degree, cm) for cm < count.  Thread-safe; each table is built once. */
int ya_ntt_init_tables(size_t degree, size_t count);
/* ntt<N>::ntt_precomputed::init(cm) on the host, into caller arrays of
 * `degree` words each; scalars[4] = {invDegree, shoupInvDegree,
 * invphiInvDegree, shoupInvphiInvDegree}.  Touches no device. */
int ya_ntt_table_host(size_t degree, size_t cm, uint64_t *phi, uint64_t *phi_shoup,
                      uint64_t *invphi, uint64_t *invphi_shoup, uint64_t *scalars);
/* Copies the device-resident table of (degree, cm) back to the host (same
 * arrays as ya_ntt_table_host), building it first if needed. */
int ya_ntt_table_download(size_t degree, size_t cm, uint64_t *phi, uint64_t *phi_shoup,
                          uint64_t *invphi, uint64_t *invphi_shoup, uint64_t *scalars);

/* ---- device memory / streams / events (plumbing) ------------------------ */
/* Replaces yell::mem_alloc / mem_free (lib/mem.cpp:14-24) for device buffers;
 * allocations are 256-byte aligned. */
int ya_malloc(void **ptr, size_t bytes);
int ya_free(void *ptr);
int ya_memset_zero(void *ptr, size_t bytes, void *stream);
int ya_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int ya_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int ya_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int ya_host_alloc_pinned(void **ptr, size_t bytes);
int ya_host_free_pinned(void *ptr);
/* *is_device = 1 if ptr is device memory of the current HIP device, else 0. */
int ya_pointer_is_device(const void *ptr, int *is_device);
int ya_stream_create(void **stream);
int ya_stream_destroy(void *stream);
int ya_stream_synchronize(void *stream);
int ya_event_create(void **event);
int ya_event_destroy(void *event);
int ya_event_record(void *event, void *stream);
int ya_event_synchronize(void *event);
int ya_event_elapsed_ms(float *ms, void *start, void *end);
/* Stream capture (no counterpart in the reference, which has no device
 * streams): every entry point called on `stream` between ya_graph_begin and
 * ya_graph_end is recorded, not run, and *graph replays the sequence with one
 * ya_graph_launch.  Call each captured entry point once beforehand (tables,
 * Gaussian tables and device state are built on first use, which capture
 * forbids).  Entry points that allocate or clear memory -- ya_memset_zero,
 * ya_sample_hwt_device, ya_poly_multiply / _many when they take the split or
 * large-ring path -- are not to be captured. */
int ya_graph_begin(void *stream);
int ya_graph_end(void *stream, void **graph);
int ya_graph_launch(void *graph, void *stream);
int ya_graph_destroy(void *graph);

/* ---- one residue (include/yell/details/ntt.hpp:106-166) ----------------- */
/* ntt<N>::forward(op, cm) / forward_lazy(op, cm) / backward(op, cm), in place
 * on `degree` device words, modulo P[cm]. */
int ya_ntt_forward(uint64_t *x, size_t degree, size_t cm, void *stream);
int ya_ntt_forward_lazy(uint64_t *x, size_t degree, size_t cm, void *stream);
int ya_ntt_backward(uint64_t *x, size_t degree, size_t cm, void *stream);

/* The out-of-line seam itself (lib/ntt.cpp:89, :151; declared at
 * include/yell/details/ntt.hpp:10-23) on HOST memory, synchronous:
 * negacylic_forward_lazy / negacylic_backward_lazy(x, degree, w, wshoup, p).
 * p must be one of P[cm]; w / wshoup must be that (degree, cm)'s phi (resp.
 * invphi) table and companions (checked on a sample of words, else
 * YA_ERR_MISMATCH).  The device uses its own resident copy of the tables. */
int ya_negacylic_forward_lazy(uint64_t *x, size_t degree, const uint64_t *w,
                              const uint64_t *wshoup, uint64_t p);
int ya_negacylic_backward_lazy(uint64_t *x, size_t degree, const uint64_t *w,
                               const uint64_t *wshoup, uint64_t p);

/* ---- residue-major polynomials (include/yell/details/poly.hpp) ---------- */
/* poly<N>::forward / forward_lazy / backward (details/poly.hpp:171-186):
 * residue r uses P[r], r < nmoduli. */
int ya_poly_forward(uint64_t *poly, size_t degree, size_t nmoduli, void *stream);
int ya_poly_forward_lazy(uint64_t *poly, size_t degree, size_t nmoduli, void *stream);
int ya_poly_backward(uint64_t *poly, size_t degree, size_t nmoduli, void *stream);
/* operator* / + / - and *= += -= (include/yell/ops.hpp:7-41, poly.hpp:163-168):
 * out[r][i] = op(a[r][i], b[r][i]).  out may alias a (the in-place forms). */
int ya_poly_mul(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree,
                size_t nmoduli, void *stream);
int ya_poly_add(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree,
                size_t nmoduli, void *stream);
int ya_poly_sub(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree,
                size_t nmoduli, void *stream);
/* poly<N>::negate (details/poly.hpp:151-158): v -> p - v (0 -> p). */
int ya_poly_negate(uint64_t *x, size_t degree, size_t nmoduli, void *stream);
/* poly<N>::add_product_of (details/poly.hpp:119-132): acc = acc + a*b mod p. */
int ya_poly_add_product_of(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t degree,
                           size_t nmoduli, void *stream);

/* The same three transforms over `count` polynomials stored back to back
 * (polynomial k at words [k*L*N, (k+1)*L*N), residue r of each using P[r]). */
int ya_poly_forward_many(uint64_t *polys, size_t degree, size_t nmoduli, size_t count, void *stream);
/* The many-polynomial forward transforms of degree 2^13 and up run residue r
 * of polynomials 2q and 2q+1 in one workgroup (shared twiddles; above 2^14,
 * after the top layers, block b of that residue); on = 0 runs one residue
 * (block) per workgroup instead (same words; default 1). */
int ya_pair_forward(int on);
int ya_poly_forward_lazy_many(uint64_t *polys, size_t degree, size_t nmoduli, size_t count,
                              void *stream);
int ya_poly_backward_many(uint64_t *polys, size_t degree, size_t nmoduli, size_t count, void *stream);
/* The pointwise operations over `count` polynomials stored back to back (as
 * the *_many transforms; per polynomial the reference's operator* / + / - /
 * add_product_of / negate, include/yell/ops.hpp:7-41,
 * include/yell/details/poly.hpp:119-132, 152-158, applied polynomial after
 * polynomial): polynomial k of out is op(a_k, b_k), with b_k = b
 * itself when b_shared != 0 (one operand for all, e.g. a key), else the k-th
 * polynomial of b.  Same words as `count` calls of the single forms; out may
 * alias a; a shared b must not overlap out.  add_product_of_many:
 * acc_k.add_product_of(a_k, b_k). */
int ya_poly_mul_many(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                     size_t count, int b_shared, void *stream);
int ya_poly_add_many(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                     size_t count, int b_shared, void *stream);
int ya_poly_sub_many(uint64_t *out, const uint64_t *a, const uint64_t *b, size_t degree, size_t nmoduli,
                     size_t count, int b_shared, void *stream);
int ya_poly_add_product_of_many(uint64_t *acc, const uint64_t *a, const uint64_t *b, size_t degree,
                                size_t nmoduli, size_t count, int b_shared, void *stream);
int ya_poly_negate_many(uint64_t *x, size_t degree, size_t nmoduli, size_t count, void *stream);
/* acc.add_product_of(a[0], b[0]); acc.add_product_of(a[1], b[1]); ... in one
 * pass (details/poly.hpp:119-132 per pair, same words): a and b are HOST
 * arrays of npairs DEVICE pointers to polys of the acc's shape. */
int ya_poly_add_products(uint64_t *acc, const uint64_t *const *a, const uint64_t *const *b,
                         size_t npairs, size_t degree, size_t nmoduli, void *stream);
/* The large-ring multiply-accumulate and its inverse in one call: for every
 * pair a_j.forward(); b_j.forward(); acc.add_product_of(a_j, b_j); then
 * acc.backward() (details/poly.hpp:119-132, 171-186), i.e. acc = backward(acc +
 * sum_j forward(a_j) * forward(b_j)) with acc given in the NTT domain
 * (canonical; zero for a plain sum of products) -- the same words as that
 * sequence.  a and b are HOST arrays of npairs DEVICE pointers to polys of
 * acc's shape; a_j and b_j are work space, their words on return unspecified
 * (from degree 2^13 both transforms of a pair and their product run in one
 * pass that writes only the product, into a_j, and the inverse's first pass
 * sums the products into acc).  The 2 * npairs + 1 buffers must not overlap. */
int ya_poly_add_products_backward(uint64_t *acc, uint64_t *const *a, uint64_t *const *b, size_t npairs,
                                  size_t degree, size_t nmoduli, void *stream);
/* The same with a fresh accumulator: out = backward(sum_j forward(a_j) *
 * forward(b_j)), npairs >= 1 -- the reference's sequence on a new (zeroed)
 * poly<N> acc(L) (details/poly.hpp:31-41 zero-fills it; details/poly.hpp:119-132,
 * 171-186), the same words as ya_poly_add_products_backward on a zeroed acc,
 * without the zero fill and without reading out.  a and b as above; out must
 * not overlap them. */
int ya_poly_sum_products_backward(uint64_t *out, uint64_t *const *a, uint64_t *const *b, size_t npairs,
                                  size_t degree, size_t nmoduli, void *stream);
/* The reference example's encryption step (example.cpp:41-44 and 46-48:
 * ct += msg; ct.forward_lazy(); ct.add_product_of(u, pk)) over `count`
 * polynomials in one pass: acc_k = forward_lazy(acc_k + add_k) + u_k * key,
 * with add null for no addition (ct_ax) and key one polynomial shared by all
 * (the public key).  Same words as ya_poly_add_many, ya_poly_forward_lazy_many
 * and ya_poly_add_product_of_many(.., b_shared = 1) in turn; fused at degrees
 * 2^13 and 2^14 (two polynomials per workgroup), those launches otherwise.
 * add, u and key must not overlap acc. */
int ya_poly_forward_lazy_add_product_many(uint64_t *acc, const uint64_t *add, const uint64_t *u,
                                          const uint64_t *key, size_t degree, size_t nmoduli, size_t count,
                                          void *stream);
/* The reference example's decryption step (example.cpp:55-57: m = ct_bx;
 * m.add_product_of(sx, ct_ax); m.backward()) over `count` polynomials:
 * out_k = backward(acc_k + x_k * key), key shared by all (the secret key).
 * Same words as a copy, ya_poly_add_product_of_many(.., b_shared = 1) and
 * ya_poly_backward_many in turn; fused up to degree 2^14.  out may be acc, or
 * else must not overlap it; out must not overlap x or key. */
int ya_poly_add_product_backward_many(uint64_t *out, const uint64_t *acc, const uint64_t *x, const uint64_t *key,
                                      size_t degree, size_t nmoduli, size_t count, void *stream);

/* One RNS poly-multiplication, fused: c = backward(forward(a) * forward(b)),
 * the call sequence a.forward(); b.forward(); c = a * b; c.backward() of the
 * reference (details/poly.hpp:171-186, ops.hpp:7-24) without modifying a, b.
 * c must not alias a or b. */
int ya_poly_multiply(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree,
                     size_t nmoduli, void *stream);

/* `count` independent poly-multiplications of the same shape in one launch:
 * polynomial k occupies words [k*L*N, (k+1)*L*N) of a, b and c, each in the
 * layout above (residue r of every polynomial uses P[r]).  Same results as
 * `count` calls of ya_poly_multiply. */
int ya_poly_multiply_many(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree,
                          size_t nmoduli, size_t count, void *stream);

/* ya_poly_multiply for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same words as ya_poly_multiply; it calls no
 * allocating routine: the device must have been set up by ya_init and the
 * tables of (degree, P[0..nmoduli)) built by ya_ntt_init_tables(degree,
 * nmoduli) beforehand (else YA_ERR_INVALID), and the scratch of the split and
 * large-ring paths is one static device array of 2^21 words (16 MiB) per
 * device, held until the process ends; a call that needs more (nmoduli *
 * degree > 2^21 above degree 2^14, or a split of more than 2^21 words)
 * returns YA_ERR_INVALID.  Every call on one device shares that array, so
 * calls that overlap on one device (two streams, two threads) race: order
 * them on one stream or serialise them. */
int ya_poly_multiply_noalloc(uint64_t *c, const uint64_t *a, const uint64_t *b, size_t degree,
                             size_t nmoduli, void *stream);

/* ---- batched form (uneven degrees and modulus counts) ------------------- *
 * No single counterpart in the reference: a batch of the reference's
 * poly-multiplications (the sequence above), possibly of different degrees
 * and modulus counts, run by one persistent kernel that drains a queue of
 * work items.  Each job's c equals what ya_poly_multiply gives for it.
 * Degrees 2^7 .. 2^15 (YA_ERR_UNSUPPORTED otherwise).  A plan fixes the
 * buffers; the caller may rewrite their contents between runs.  Runs of one
 * plan must not overlap (one stream at a time). */
typedef struct ya_polymul_job {
  const uint64_t *a;  /* device, [nmoduli][degree] */
  const uint64_t *b;
  uint64_t *c;        /* must not overlap a or b */
  uint32_t degree;
  uint32_t nmoduli;
} ya_polymul_job;
typedef struct ya_batch_plan ya_batch_plan;
int ya_batch_plan_create(ya_batch_plan **plan, const ya_polymul_job *jobs, size_t njobs);
/* Launches the plan's kernels on `stream`. */
int ya_batch_plan_run(ya_batch_plan *plan, void *stream);
/* Same work as a hipGraph: captured on first use, then replayed on `stream`. */
int ya_batch_plan_run_graph(ya_batch_plan *plan, void *stream);
/* Launch forms: a batch of one degree class <= 2^14 holding at least two
 * residues per CU runs as a grid of the fused kernel over the plan's residue
 * table (YA_BATCH_GRID); a batch of one class 2^11 .. 2^14 too small to
 * occupy half of the CUs spreads each residue over several workgroups
 * (YA_BATCH_SPLIT, see ya_split_policy; the plan holds a scratch copy of the
 * b operands' partial transforms); any other batch drains the persistent
 * kernel's queue (YA_BATCH_PERSISTENT).  All give the same words. */
enum { YA_BATCH_PERSISTENT = 1, YA_BATCH_GRID = 2, YA_BATCH_SPLIT = 3 };
int ya_batch_plan_info(const ya_batch_plan *plan, size_t *items, size_t *residues,
                       size_t *workgroups, int *form);
/* Forces the launch form of plans created afterwards (0 = automatic; GRID
 * and SPLIT only where the batch allows them). */
int ya_batch_force_form(int form);
/* Plans created while on (default) run a grid-form batch of back-to-back
 * polynomials of one shape (job k at a_0 + k*N*L, likewise b and c, c apart
 * from a and b) with the fused kernel addressing residues directly instead of
 * through the residue table; same words.  0: always the residue table. */
int ya_batch_strided(int on);
/* *strided = 1 if the plan runs the direct-addressed path described above. */
int ya_batch_plan_is_strided(const ya_batch_plan *plan, int *strided);
int ya_batch_plan_destroy(ya_batch_plan *plan);
/* One-shot: create, run on `stream`, wait, destroy. */
int ya_batch_polymul(const ya_polymul_job *jobs, size_t njobs, void *stream);

/* Autotuner (no counterpart in the reference).  For every degree class
 * 2^lo .. 2^hi (within 2^7 .. 2^15) times, on the device, the instantiated
 * shapes of the batched (persistent) kernel -- radix-8 or radix-16 register
 * passes, i.e. 3 or 4 NTT layers fused per pass; the radix also fixes the
 * tile: how many residues share a 1024-thread workgroup -- and of the fused
 * single-launch kernel -- radix 8, 16 or 32 (one residue per workgroup of
 * 2^(logn - log2 radix) threads), and from 2^10 at radix 8 / 16 both forward
 * transforms as one pass sequence with shared twiddles (dual) or as two --
 * and keeps the fastest of each for later plans / launches (table kinds 2,
 * 1 and 3; kind 0 follows kind 1).  report (optional) receives
 * YA_TUNE_REPORT_STRIDE doubles per class: batched ms at radix 8, 16; fused
 * ms at radix 8, 16 (dual where built), 32; fused ms at radix 8, 16 as two
 * sequences (-1 where not applicable); each is the best of 5 launches taken
 * round-robin over the class's shapes, after ~0.15 s of untimed launches
 * that bring the GPU's clocks up.  Results are identical for every shape. */
enum { YA_TUNE_REPORT_STRIDE = 7 };
int ya_autotune(int logn_lo, int logn_hi, double *report);
/* Read / set one entry of the shape table: kind 1 = fused kernel, 2 =
 * batched kernel, 0 = standalone NTT kernels; le = log2 of the register-pass
 * radix (3, 4, or 5 for kinds 0 and 1).  kind 3: the fused kernel's
 * dual-forward form (both operands' forward transforms in one pass sequence
 * with shared twiddles; radix 8 / 16 at 2^10 and up), value 0 or 1.  kind 4:
 * the block size (log2: 14, 13 up to degree 2^18, or 12 up to 2^17; default
 * 12 from 2^15 to 2^17, 13 at 2^18, 14 otherwise) of the pair-product pass
 * of ya_poly_add_products_backward / ya_poly_sum_products_backward.  kind 5:
 * staggered start, 0..16 steps of ~0.85 us (0 = off), of the fused kernel at
 * degree 2^logn and, at logn = 15,
 * of the batched (persistent) kernel: in a launch of at least eight waves of
 * workgroups, the first wave's workgroups wait (group mod 16) steps before
 * they start, so the CUs' operand loads do not all fall together; default 2
 * at 2^14, 6 at 15, 0 elsewhere.  No kind changes a result.
 * ya_tune_get returns the entry's value, or -1. */
int ya_tune_get(int kind, int logn);
int ya_tune_set(int kind, int logn, int le);
/* Autotuner of the large-ring multiply-accumulate (no counterpart in the
 * reference): times ya_poly_sum_products_backward at this degree (above
 * 2^14), nmoduli (up to 65535) and npairs (1..16) on zero operands in its
 * own device buffers, with each accepted pair-pass block size, and keeps the
 * fastest as tune kind 4 for the degree.  report (optional) receives 3
 * doubles: ms per job with 2^14, 2^13, 2^12 blocks (-1 where not accepted).
 * It rewrites kind 4 while it runs: not concurrently with other calls at
 * that degree. */
int ya_autotune_mac(size_t degree, size_t nmoduli, size_t npairs, double *report);
/* Split launches (no counterpart in the reference).  A poly-multiplication
 * launch of few residues of degree 2^11 .. 2^14 (ya_poly_multiply(_many), the
 * served form's small batches) may spread each residue over 2^k workgroups
 * (column kernels for the top k index bits, block kernels for the rest)
 * instead of one, for latency.  mode 0: automatic (when one workgroup per
 * residue would occupy at most half of the CUs); -1: never; k in 2..5:
 * always, with that k, where 2^logn / 2^k >= 2^9.  Results are identical. */
int ya_split_policy(int mode);

/* ---- served form (request loop with continuous batching) --------------- *
 * No counterpart in the reference.  Clients submit poly-multiplication
 * requests with HOST operands a, b and a host result buffer c (pinned memory
 * from ya_host_alloc_pinned gives asynchronous copies); the server admits
 * pending requests while their device footprint (3 * degree * nmoduli words)
 * fits in an arena of hbm_budget_bytes, batches whatever is admitted into one
 * launch of the batched kernel on one of max_inflight streams, and retires
 * batches as they finish.  Latency = submit -> result in c.  Results equal
 * ya_poly_multiply's.  Buffers must stay valid until the request is waited
 * for.  Degrees 2^7 .. 2^15, nmoduli <= max_moduli. */
typedef struct ya_server ya_server;
typedef struct ya_server_config {
  size_t hbm_budget_bytes;  /* device arena for in-flight operands (default 1 GiB) */
  uint32_t max_batch_jobs;  /* requests per launch (default 256) */
  uint32_t max_inflight;    /* batches on the device at once, <= 8 (default 2) */
  uint32_t max_pending;     /* queued requests before submit is rejected (default 4096) */
  uint32_t max_moduli;      /* largest nmoduli accepted (default 16) */
} ya_server_config;          /* zero fields take the defaults */
typedef struct ya_server_stats {
  uint64_t submitted, completed, rejected, batches, max_pending, arena_bytes_in_use;
  double mean_batch_jobs, p50_us, p90_us, p99_us, max_us, mean_us;
  /* split of the latency: queue = submit -> dispatched in a batch,
   * service = dispatched -> result in c (copies + kernel + retirement) */
  double p50_queue_us, p99_queue_us, p50_service_us, p99_service_us;
  /* per batch: device time of copies-in + kernel, device time of copies-out,
   * host time issuing the batch, host time dispatch -> retired */
  double p50_batch_gpu_us, p99_batch_gpu_us, p99_batch_d2h_us, p99_batch_launch_us;
  double p50_batch_host_us, p99_batch_host_us;
} ya_server_stats;
int ya_server_create(ya_server **server, const ya_server_config *config);
/* YA_ERR_REJECTED when the request can never fit or the queue is full. */
int ya_server_submit(ya_server *server, const uint64_t *a, const uint64_t *b, uint64_t *c,
                     uint32_t degree, uint32_t nmoduli, uint64_t *ticket);
/* Blocks until the request is done (timeout_ms < 0: no limit); releases the ticket. */
int ya_server_wait(ya_server *server, uint64_t ticket, int timeout_ms, double *latency_us);
int ya_server_poll(ya_server *server, uint64_t ticket, int *done);
int ya_server_get_stats(ya_server *server, ya_server_stats *stats);
int ya_server_reset_stats(ya_server *server);
/* Stops the loop; requests not yet dispatched fail. */
int ya_server_destroy(ya_server *server);

/* Poisson load generator over a server: exponential inter-arrival gaps at
 * rate_per_s for duration_s; each request picks one of nshapes (degree,
 * nmoduli) shapes uniformly and one of `pool` seeded operand pairs of that
 * shape (a uniform, b Gaussian sigma 3.2).  verify != 0 compares every served
 * result with the same pair multiplied alone before the run. */
typedef struct ya_poisson_spec {
  double rate_per_s, duration_s;
  uint32_t nshapes;
  const uint32_t *degrees, *nmoduli;
  uint32_t pool, verify;
  uint64_t seed;
} ya_poisson_spec;
typedef struct ya_serve_report {
  uint64_t offered, completed, rejected, failed, mismatches, batches, max_pending;
  uint64_t client_stalls;  /* arrivals delayed because every result buffer was in use */
  double wall_s, throughput_per_s, p50_us, p90_us, p99_us, max_us, mean_us, mean_batch_jobs;
  double p50_queue_us, p99_queue_us, p50_service_us, p99_service_us;
  double p50_batch_gpu_us, p99_batch_gpu_us, p99_batch_d2h_us, p99_batch_launch_us;
  double p50_batch_host_us, p99_batch_host_us;
} ya_serve_report;
int ya_server_run_poisson(ya_server *server, const ya_poisson_spec *spec, ya_serve_report *report);

/* ---- host samplers (input source; details/poly.hpp:200-330) ------------- */
/* Seeded replacements of poly::set(uniform) and poly::set(gaussian): fill a
 * HOST residue-major buffer.  The random bytes are the stream of `seed`:
 * block b (32 bytes, four little-endian 64-bit words) is Philox4x64-10 with
 * counter (b, 0, 0, 0) and key (seed, 0).  uniform (details/poly.hpp:200-228):
 * residue cm takes stream words cm*degree .. cm*degree+degree-1, each masked to
 * the prime's bit length (2^62-1) and reduced by p once if >= p.  gaussian
 * (details/poly.hpp:236-260): one centred discrete Gaussian sample per
 * coefficient (density proportional to exp(-x^2/(2 sigma^2)), sigma in
 * (0, 1e6]), drawn from the table of ya_gaussian_cdt with stream word 2i
 * (magnitude) and the top bit of word 2i+1 (sign), shared by all residues and
 * mapped r < 0 ? p + r : r. */
int ya_sample_uniform(uint64_t *out, size_t degree, size_t nmoduli, uint64_t seed);
int ya_sample_gaussian(uint64_t *out, size_t degree, size_t nmoduli, double sigma, uint64_t seed);
/* poly::set(ZO_dist) (details/poly.hpp:267-279): stream byte i is
 * coefficient i's r, r <= rho ? ((r & 2) ? p-1 : 1) : 0, shared by all
 * residues. */
int ya_sample_zo(uint64_t *out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed);
/* poly::set(hwt_dist) (details/poly.hpp:287-330): exactly hwt nonzero
 * coefficients, shared by all residues.  Reservoir: slots 0..hwt-1 start at
 * positions 0..hwt-1; step k = hwt..degree-1 draws pos uniform below k
 * (first of the four words of Philox block (k, 1, 0, 0) with
 * w <= floor((2^64-1)/k)*k, else the fourth; pos = w mod k) and replaces
 * slot pos if pos < hwt.  The j-th smallest hit position is +1 if stream
 * word j has bit 1 set, else p-1. */
int ya_sample_hwt(uint64_t *out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed);
/* FastGaussianNoise::getNoise (include/yell/prng/FastGaussianNoise.hpp:484-602)
 * over this library's stream: out[i] = the signed sample of coefficient i of
 * the stream of `seed` (the values ya_sample_gaussian maps to residues), for
 * any count.  The C++ drop-in's FastGaussianNoise<In, Out, D> draws from it. */
int ya_gaussian_noise(int64_t *out, size_t count, double sigma, uint64_t seed);
/* The Gaussian samplers' table: cdf[k] = floor(P(|X| <= k) * 2^64) for
 * k < tail = ceil(13 sigma), cdf[tail] = 2^64-1; *count = tail + 1.  out may
 * be NULL to query the count. */
int ya_gaussian_cdt(double sigma, uint64_t *out, size_t capacity, size_t *count);

/* ---- device samplers (details/poly.hpp:200-330 on the device) ------------
 * The same streams and mappings as the host samplers above, written straight
 * into a DEVICE residue-major buffer (16-byte aligned) on `stream`: the call
 * with the same seed returns the same words as ya_sample_*.  No counterpart in
 * the reference (its samplers are host-only); they replace host sampling plus
 * the upload.  gaussian: the signed sample r is multiplied by `amplifier`
 * (gaussian::amplifier, details/poly.hpp:248-251) in the reference's
 * params::signed_type -- int64_t here, int32_t in the ya32 build
 * (params_32.hpp), i.e. the product wraps modulo 2^64 (2^32) -- and mapped
 * a < 0 ? p + a : a modulo 2^64 (2^32); amplifier 1 is ya_sample_gaussian.  hwt uses
 * stream-ordered scratch (hipMallocAsync) of hwt + 2*ceil(degree/32) words. */
int ya_sample_uniform_device(uint64_t *d_out, size_t degree, size_t nmoduli, uint64_t seed, void *stream);
int ya_sample_gaussian_device(uint64_t *d_out, size_t degree, size_t nmoduli, double sigma, uint64_t amplifier,
                              uint64_t seed, void *stream);
int ya_sample_zo_device(uint64_t *d_out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed, void *stream);
int ya_sample_hwt_device(uint64_t *d_out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed, void *stream);
/* `count` polynomials back to back in one launch: polynomial k (words
 * [k*L*N, (k+1)*L*N)) holds what the single call with seed + k writes. */
int ya_sample_uniform_device_many(uint64_t *d_out, size_t degree, size_t nmoduli, size_t count, uint64_t seed,
                                  void *stream);
int ya_sample_gaussian_device_many(uint64_t *d_out, size_t degree, size_t nmoduli, size_t count, double sigma,
                                   uint64_t amplifier, uint64_t seed, void *stream);
int ya_sample_zo_device_many(uint64_t *d_out, size_t degree, size_t nmoduli, size_t count, uint32_t rho,
                             uint64_t seed, void *stream);

/* ---- CRT lift: residues -> multi-limb integers ----------------------------
 * Replaces poly<N>::poly2mpz (include/yell/details/poly.hpp:162-165), i.e.
 * gmp::poly2mpz (include/yell/gmp.hpp:41-70) with the constants of
 * gmp::gmp_precomputed (lib/gmp.cpp:6-51), for the primes P[0..nmoduli).
 * Each coefficient i becomes S mod Q, Q = P[0]*...*P[nmoduli-1], where
 * S = sum_j lift_j * r_j and lift_j = (Q/P[j]) * ((Q/P[j])^-1 mod P[j]) -- the
 * value GMP returns there for any 64-bit residues r_j (canonical or not).
 *
 * ya_crt_limbs: 64-bit limbs of a lifted coefficient, ceil(bits(Q) / 64).
 * ya_poly_crt_lift: residues (device, [nmoduli][degree]) -> out (device,
 *   [degree][limbs], little-endian limbs per coefficient: the layout
 *   mpz_import(rop, limbs, -1, 8, 0, 0, out + i * limbs) reads); any degree
 *   in [1, 2^20]; stream-ordered.  out must not overlap residues.
 * ya_crt_constants: the host copies of moduli_product_ (limbs limbs),
 *   modulus_shoup (2 limbs), lifting_integers ([nmoduli][limbs]) and
 *   shift_modulus_shoup; any output pointer may be NULL. */
int ya_crt_limbs(size_t nmoduli, size_t *limbs);
int ya_poly_crt_lift(uint64_t *out, const uint64_t *residues, size_t degree, size_t nmoduli,
                     void *stream);
int ya_crt_constants(size_t nmoduli, size_t limbs, uint64_t *moduli_product,
                     uint64_t *modulus_shoup, uint64_t *lifting, size_t *shift);

/* ---- 32-bit moduli build -------------------------------------------------
 * The reference's compile-time alternative (CMakeLists.txt:19
 * YELL_USE_32BITS_MODULI -> include/yell/params_32.hpp): uint32_t words,
 * uint64_t double words, 291 primes 2^30 - k*2^16 + 1 (kModulusBitsize 30),
 * degrees 2^3 .. 2^15 (kMaxPolyDegree 2^15).  Same semantics, layout and value
 * ranges as the 64-bit entry points of the same name without the 32 (the
 * reference code is templated on params::value_type; every butterfly, Barrett
 * and Shoup step is the same with 32-bit words): ntt<N>::forward /
 * forward_lazy / backward (details/ntt.hpp:106-166), poly operators
 * (ops.hpp:7-41, details/poly.hpp:119-186), and the fused poly-multiplication.
 * Device buffers hold uint32_t words. */
size_t ya32_num_moduli(void);
uint32_t ya32_prime(size_t cm);
uint32_t ya32_barrett_const(size_t cm);
uint32_t ya32_root(size_t cm);
int ya32_ntt_table_host(size_t degree, size_t cm, uint32_t *phi, uint32_t *phi_shoup,
                        uint32_t *invphi, uint32_t *invphi_shoup, uint32_t *scalars);
int ya32_ntt_init_tables(size_t degree, size_t count);
int ya32_ntt_forward(uint32_t *x, size_t degree, size_t cm, void *stream);
int ya32_ntt_forward_lazy(uint32_t *x, size_t degree, size_t cm, void *stream);
int ya32_ntt_backward(uint32_t *x, size_t degree, size_t cm, void *stream);
int ya32_poly_forward(uint32_t *poly, size_t degree, size_t nmoduli, void *stream);
int ya32_poly_forward_lazy(uint32_t *poly, size_t degree, size_t nmoduli, void *stream);
int ya32_poly_backward(uint32_t *poly, size_t degree, size_t nmoduli, void *stream);
int ya32_poly_mul(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree,
                  size_t nmoduli, void *stream);
int ya32_poly_add(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree,
                  size_t nmoduli, void *stream);
int ya32_poly_sub(uint32_t *out, const uint32_t *a, const uint32_t *b, size_t degree,
                  size_t nmoduli, void *stream);
int ya32_poly_negate(uint32_t *x, size_t degree, size_t nmoduli, void *stream);
int ya32_poly_add_product_of(uint32_t *acc, const uint32_t *a, const uint32_t *b, size_t degree,
                             size_t nmoduli, void *stream);
int ya32_poly_multiply(uint32_t *c, const uint32_t *a, const uint32_t *b, size_t degree,
                       size_t nmoduli, void *stream);
int ya32_poly_multiply_many(uint32_t *c, const uint32_t *a, const uint32_t *b, size_t degree,
                            size_t nmoduli, size_t count, void *stream);
/* Host samplers of the 32-bit build (same rules as ya_sample_*, with the
 * 30-bit primes; uniform reads the stream as 32-bit words, as the 32-bit
 * reference fills its uint32 residues with random bytes: residue cm takes
 * 32-bit words cm*degree .. cm*degree+degree-1, masked to 30 bits, minus p
 * once). */
int ya32_sample_uniform(uint32_t *out, size_t degree, size_t nmoduli, uint64_t seed);
int ya32_sample_gaussian(uint32_t *out, size_t degree, size_t nmoduli, double sigma, uint64_t seed);
int ya32_sample_zo(uint32_t *out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed);
int ya32_sample_hwt(uint32_t *out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed);
/* Device samplers of the 32-bit build (as ya_sample_*_device; uniform takes
 * 32-bit stream words, residue cm from word cm*degree on). */
int ya32_sample_uniform_device(uint32_t *d_out, size_t degree, size_t nmoduli, uint64_t seed, void *stream);
int ya32_sample_gaussian_device(uint32_t *d_out, size_t degree, size_t nmoduli, double sigma, uint64_t amplifier,
                                uint64_t seed, void *stream);
int ya32_sample_zo_device(uint32_t *d_out, size_t degree, size_t nmoduli, uint32_t rho, uint64_t seed, void *stream);
int ya32_sample_hwt_device(uint32_t *d_out, size_t degree, size_t nmoduli, uint32_t hwt, uint64_t seed, void *stream);

/* ---- measurement ---------------------------------------------------------
 * No counterpart in the reference.  Device throughput of the VALU work this
 * engine is bound by, all operands in registers: kind 0 = v_mad_u64_u32
 * (32x32+64 -> 64), kind 1 = v_mul_lo_u32, kind 2 = this engine's lazy CT
 * butterfly (lib/ntt.cpp:69-83), kind 3 = the 32-bit moduli build's CT
 * butterfly.  Result in operations (butterflies) per second over the whole
 * device. */
int ya_probe_valu(int kind, double *ops_per_second);

#ifdef __cplusplus
}
#endif
#endif /* YELL_AMD_H */
