#!/usr/bin/env python3
"""Throughput bench of the RNS poly-multiplication path on one MI355X.

Workload: N = 2^14, L = 8 (primes P[0..7]); operand a uniform per residue,
operand b a centred discrete Gaussian (sigma = 3.2) reduced mod each prime,
both from fixed seeds; one poly-mult = forward(a), forward(b), Barrett
pointwise product, backward.  A step is one batched launch over --batch
polynomials already resident in HBM.  The same polynomials are also run as
plain one-poly-per-launch calls, and a bounded sample is timed on the host
CPU with the reference's own code (oracle/_ref, compiled from its sources) or,
where that is absent, the oracle restatement.

Prints ONE JSON line.  `value` = batched poly-mults/s.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from yell_amd import capi  # noqa: E402

METRIC = "bit-exact poly-mults/s at N=2^14,L=8 batched vs plain; p50/p99 served latency"
HBM_PEAK_GBS = 8000.0
# HBM bytes one poly-multiplication of the batched step moves, from the
# FETCH_SIZE / WRITE_SIZE counter passes (calibrated against a 256 MiB copy)
# in profiles/k_polymul_dual_n16384_l8.md, per launch of 512 poly-mults:
# k_polymul_loop (the direct-addressed looping kernel a plan over 512
# back-to-back polynomials runs) 1.1365e9 B read + 5.369e8 B written;
# k_polymul (one workgroup per residue) 1.1443e9 + 5.369e8; k_polymul_tab
# (the residue-table kernel) 1.1535e9 + 5.369e8.
TRAFFIC_BYTES_PER_POLYMUL = {"k_polymul_loop<14,4>": 3268262.0, "k_polymul<14,4,dual>": 3283596.0,
                             "k_polymul_tab<14,4,dual>": 3301576.0}
# The binding resource of the fused kernel is VALU issue: one wave64 VALU
# instruction per CU per cycle (4 SIMDs of 16 lanes, 4 cycles per wave64
# instruction), 256 CUs at the 2.4 GHz peak engine clock.  VALU wave-
# instructions per poly-multiplication from SQ_INSTS_VALU at --batch 512
# (profiles/k_polymul_dual_n16384_l8.md): the looping kernel 525,484,032 per
# launch = 128,292 per residue x 8 residues; one workgroup per residue
# 8,097 per wave x 16 waves per residue x 8 residues.
CUS, ENGINE_CLOCK_GHZ = 256, 2.4
VALU_PEAK_G_PER_S = CUS * ENGINE_CLOCK_GHZ
VALU_INSTS_PER_POLYMUL = {"k_polymul_loop<14,4>": 128292 * 8, "k_polymul<14,4,dual>": 8097 * 16 * 8}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def modmuls_per_polymul(n, L):
    lg = n.bit_length() - 1
    fwd = (n // 2) * lg
    inv = (n // 2) * (lg - 1) + n
    return L * (2 * fwd + inv + n)


class Events:
    def __init__(self):
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        capi.call("ya_event_create", ctypes.byref(self.a))
        capi.call("ya_event_create", ctypes.byref(self.b))

    def elapsed_ms(self):
        capi.call("ya_event_synchronize", self.b)
        ms = ctypes.c_float()
        capi.call("ya_event_elapsed_ms", ctypes.byref(ms), self.a, self.b)
        return ms.value


def make_inputs(n, L, batch, seed):
    a = np.empty(batch * n * L, dtype=np.uint64)
    b = np.empty(batch * n * L, dtype=np.uint64)
    w = n * L
    for k in range(batch):
        a[k * w:(k + 1) * w] = capi.sample_uniform(n, L, seed + 2 * k)
        b[k * w:(k + 1) * w] = capi.sample_gaussian(n, L, 3.2, seed + 2 * k + 1)
    return a, b


def cpu_baseline(a, b, n, L, gpu_c, sample_polys, seconds):
    """Time the reference's poly-mult on the host cores (bounded sample) and
    check the GPU results of polynomials `sample_polys` against it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle  # checker only: this leg measures and verifies, it is never the product

    w = n * L
    a0, b0 = a[:w].copy(), b[:w].copy()
    threads = max(1, min(16, os.cpu_count() or 1))
    try:
        ref = _oracle.RefLib()
        kind = "reference"
        polymul = ref.polymul
        lg = n.bit_length() - 1
        t1 = ref.lib.yref_bench_polymul(lg, capi.ptr(a0), capi.ptr(b0), L, 1, 1)
        reps = max(1, int(seconds / max(t1, 1e-6) / 1.0))
        reps_1 = max(1, int(0.25 * seconds / max(t1, 1e-6)))
        s1 = ref.lib.yref_bench_polymul(lg, capi.ptr(a0), capi.ptr(b0), L, 1, reps_1)
        sN = ref.lib.yref_bench_polymul(lg, capi.ptr(a0), capi.ptr(b0), L, threads,
                                        max(1, reps // threads * 1))
        per_thread = max(1, reps // threads)
        value = threads * per_thread / sN
        single = reps_1 / s1
        sample = ("%d threads x %d poly-mults (N=%d, L=%d) of the reference's own ntt<N>/ops code, "
                  "Release flags; single core: %d poly-mults" % (threads, per_thread, n, L, reps_1))
    except RuntimeError:
        orc = _oracle.Oracle()
        kind = "port"
        threads = 1
        polymul = orc.polymul
        t0 = time.perf_counter()
        orc.polymul(a0, b0, n, L)
        t1 = time.perf_counter() - t0
        reps = max(1, int(seconds / max(t1, 1e-6)))
        t0 = time.perf_counter()
        for _ in range(reps):
            orc.polymul(a0, b0, n, L)
        value = reps / (time.perf_counter() - t0)
        single = value
        sample = "%d poly-mults (N=%d, L=%d) of the C restatement, one thread" % (reps, n, L)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    exact = all(np.array_equal(polymul(a[k * w:(k + 1) * w].copy(), b[k * w:(k + 1) * w].copy(), n, L),
                               gpu_c[k * w:(k + 1) * w]) for k in sample_polys)
    return {"value": value, "unit": "poly-mults/s", "cores": threads, "kind": kind,
            "sample": sample, "single_core_value": single, "cpu_model": model,
            "host_cpus": os.cpu_count()}, exact


def roofline(kernel, ms, polys, hbm_achieved, bytes_per):
    """The step kernel against the resource that binds it (VALU issue), with
    the HBM side as a secondary block."""
    traffic = TRAFFIC_BYTES_PER_POLYMUL[kernel] * polys if kernel in TRAFFIC_BYTES_PER_POLYMUL else None
    hbm = {"achieved": hbm_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm_achieved / HBM_PEAK_GBS,
           "bytes_per_polymul": bytes_per}
    common = {"traffic": traffic, "traffic_source": "profiles/k_polymul_dual_n16384_l8.md", "kernel": kernel,
              "polymuls_per_launch": polys}
    if kernel not in VALU_INSTS_PER_POLYMUL:
        return dict({"bound": "hbm"}, **hbm, **common)
    insts = VALU_INSTS_PER_POLYMUL[kernel] * polys
    valu = insts / (ms / 1e3) / 1e9
    return dict({"bound": "valu", "achieved": valu, "peak": VALU_PEAK_G_PER_S,
                 "unit": "G VALU wave-instructions/s", "frac": valu / VALU_PEAK_G_PER_S,
                 "valu_insts_per_launch": insts,
                 "peak_basis": "%d CUs x %.1f GHz x 1 wave64 VALU instruction per CU-cycle" % (CUS, ENGINE_CLOCK_GHZ),
                 "hbm": hbm}, **common)


def make_plan(jobs_spec):
    """jobs_spec: list of (a_ptr, b_ptr, c_ptr, n, L) -> plan handle."""
    arr = (capi.PolymulJob * len(jobs_spec))()
    for k, (a, b, c, n, L) in enumerate(jobs_spec):
        arr[k] = capi.PolymulJob(a, b, c, n, L)
    plan = ctypes.c_void_p()
    capi.call("ya_batch_plan_create", ctypes.byref(plan), ctypes.cast(arr, ctypes.c_void_p),
              len(jobs_spec))
    return plan


def timed(ev, stream, fn, steps, settle_s=0.1):
    """Mean ms of `steps` calls of fn, timed with events on `stream`, after
    `settle_s` seconds of untimed calls: after host-side work (input
    generation, uploads, downloads) the GPU has idled and its clocks take
    ~0.1 s of work to ramp back up."""
    t0 = time.perf_counter()
    while True:
        fn()
        capi.call("ya_stream_synchronize", stream)
        if time.perf_counter() - t0 >= settle_s:
            break
    capi.call("ya_event_record", ev.a, stream)
    for _ in range(steps):
        fn()
    capi.call("ya_event_record", ev.b, stream)
    return ev.elapsed_ms() / steps


def c3_bench(stream, ev, steps, seed=2026):
    """1024 uneven jobs (degrees 2^10..2^15, 1..12 moduli) in one persistent
    launch (hipGraph), against one ya_poly_multiply launch per job."""
    rng = np.random.default_rng(seed)
    specs = [(1 << int(rng.integers(10, 16)), int(rng.integers(1, 13))) for _ in range(1024)]
    total = sum(n * L for n, L in specs)
    A, B, C = (capi.DeviceBuffer(total) for _ in range(3))
    host = np.empty(total, dtype=np.uint64)
    off, jobs = 0, []
    for k, (n, L) in enumerate(specs):
        host[off:off + n * L] = capi.sample_uniform(n, L, seed + k)
        jobs.append((A.ptr + 8 * off, B.ptr + 8 * off, C.ptr + 8 * off, n, L))
        off += n * L
    A.upload(host)
    B.upload(host[::-1].copy())
    for n in sorted({n for n, _ in specs}):
        capi.call("ya_ntt_init_tables", n, 12)
    plan = make_plan(jobs)
    run = lambda: capi.call("ya_batch_plan_run_graph", plan, stream)
    run()
    capi.call("ya_stream_synchronize", stream)
    ms = timed(ev, stream, run, steps)
    plain = lambda: [capi.call("ya_poly_multiply", c, a, b, n, L, stream) for a, b, c, n, L in jobs]
    plain()
    ms_plain = timed(ev, stream, plain, 1)
    capi.load().ya_batch_plan_destroy(plan)
    for d in (A, B, C):
        d.close()
    return {"jobs": len(specs), "words_per_operand": total,
            "value": len(specs) / (ms / 1e3), "unit": "poly-mults/s", "ms_per_batch": ms,
            "plain_value": len(specs) / (ms_plain / 1e3), "ms_plain": ms_plain,
            "batched_over_plain": ms_plain / ms,
            "modmuls_T_per_s": sum(modmuls_per_polymul(n, L) for n, L in specs) / (ms / 1e3) / 1e12,
            "hbm_GBps_algorithmic": 3 * 8 * total / (ms / 1e3) / 1e9}


def c4_bench(stream, ev, steps):
    """N = 2^16, L = 24: forward NTT of 16 operands, sum of 8 pair products in
    the NTT domain, one inverse NTT (17 * N * L * 8 algorithmic bytes).  The
    job is ya_poly_sum_products_backward (a fresh accumulator: the pair-product
    pass and the summing inverse, which writes the result without reading it;
    the pair pass's block size chosen by ya_autotune_mac, which times 2^14,
    2^13 and 2^12 blocks); beside it
    ya_poly_add_products_backward after a zero fill of acc (the same words,
    checked) and the reference's sequence as separate launches (forward of the
    16 operands, the MAC, the inverse)."""
    n, L, pairs = 1 << 16, 24, 8
    w = n * L
    ops = np.concatenate([capi.sample_uniform(n, L, 4000 + k) for k in range(2 * pairs)])
    dall, dacc = capi.DeviceBuffer.from_host(ops), capi.DeviceBuffer(w)
    capi.call("ya_ntt_init_tables", n, L)
    pa = (ctypes.c_void_p * pairs)(*[dall.ptr + 8 * w * (2 * j) for j in range(pairs)])
    pb = (ctypes.c_void_p * pairs)(*[dall.ptr + 8 * w * (2 * j + 1) for j in range(pairs)])
    e2 = Events()

    def fused():
        capi.call("ya_memset_zero", dacc.ptr, w * 8, stream)
        capi.call("ya_poly_add_products_backward", dacc.ptr, pa, pb, pairs, n, L, stream)

    def fresh():
        capi.call("ya_poly_sum_products_backward", dacc.ptr, pa, pb, pairs, n, L, stream)

    def separate():
        capi.call("ya_poly_forward_many", dall.ptr, n, L, 2 * pairs, stream)
        capi.call("ya_event_record", e2.a, stream)
        capi.call("ya_memset_zero", dacc.ptr, w * 8, stream)
        capi.call("ya_poly_add_products", dacc.ptr, pa, pb, pairs, n, L, stream)
        capi.call("ya_event_record", e2.b, stream)
        capi.call("ya_poly_backward", dacc.ptr, n, L, stream)

    lg = n.bit_length() - 1
    # the library's tuner picks the pair pass's block size for this shape
    rep = (ctypes.c_double * 3)()
    capi.call("ya_autotune_mac", n, L, pairs, rep)
    tuned = {str(14 - k): round(rep[k], 4) for k in range(3) if rep[k] > 0}
    best = capi.load().ya_tune_get(4, lg)
    fresh()
    capi.call("ya_stream_synchronize", stream)
    ms = timed(ev, stream, fresh, steps)
    fused()
    capi.call("ya_stream_synchronize", stream)
    ms_acc = timed(ev, stream, fused, steps)
    # the two forms give the same words (each call uses a_j, b_j as work
    # space, so the operands are uploaded again before each)
    dall.upload(ops)
    fresh()
    capi.call("ya_stream_synchronize", stream)
    got_fresh = dacc.download()
    dall.upload(ops)
    fused()
    capi.call("ya_stream_synchronize", stream)
    same = bool(np.array_equal(got_fresh, dacc.download()))
    separate()
    capi.call("ya_stream_synchronize", stream)
    ms_sep = timed(ev, stream, separate, steps)
    ms_mac = e2.elapsed_ms()
    dall.close()
    dacc.close()
    alg = 17 * w * 8
    big = {}
    for lgn, Lb in ((16, 24), (19, 4)):  # one large-ring poly-multiplication (ya_poly_multiply)
        nb = 1 << lgn
        a1 = capi.DeviceBuffer.from_host(capi.sample_uniform(nb, Lb, 77))
        b1 = capi.DeviceBuffer.from_host(capi.sample_gaussian(nb, Lb, 3.2, 78))
        c1 = capi.DeviceBuffer(nb * Lb)
        one = lambda: capi.call("ya_poly_multiply", c1.ptr, a1.ptr, b1.ptr, nb, Lb, stream)
        one()
        capi.call("ya_stream_synchronize", stream)
        ms1 = timed(ev, stream, one, steps)
        big["n%d_l%d" % (nb, Lb)] = {"ms": ms1, "algorithmic_bytes": 3 * nb * Lb * 8,
                                     "hbm_GBps_algorithmic": 3 * nb * Lb * 8 / (ms1 / 1e3) / 1e9}
        for d_ in (a1, b1, c1):
            d_.close()
    return {"degree": n, "nmoduli": L, "pairs": pairs, "ms_per_job": ms, "large_ring_polymul": big,
            "jobs_per_s": 1e3 / ms, "algorithmic_bytes": alg,
            "hbm_GBps_algorithmic": alg / (ms / 1e3) / 1e9,
            "entry_point": "ya_poly_sum_products_backward",
            "ms_by_pair_block_log_tuner": tuned, "pair_block_log": best,
            "accumulate_form": {"entry_point": "ya_poly_add_products_backward (after a zero fill of acc)",
                                "ms_per_job": ms_acc, "same_words_as_fresh": same},
            "separate_launches": {"ms_per_job": ms_sep, "mac_ms": ms_mac,
                                  "mac_GBps": (2 * pairs + 2) * w * 8 / (ms_mac / 1e3) / 1e9}}


def u32_bench(stream, ev, reps=5):
    """32-bit moduli build (ya32_poly_multiply_many): N = 2^14 with 16 primes
    of 30 bits (a ~480-bit modulus, close to the 8 x 62-bit modulus of the headline), 256
    polys per launch (4096 residues, as the headline step)."""
    n, L, count = 1 << 14, 16, 256
    w = n * L
    rng = np.random.default_rng(5)
    primes = [int(capi.load().ya32_prime(j)) for j in range(L)]
    a = np.concatenate([rng.integers(0, primes[j], size=n, dtype=np.uint32) for j in range(L)] * count)
    b = np.concatenate([rng.integers(0, primes[j], size=n, dtype=np.uint32) for j in range(L)] * count)
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    dc = capi.DeviceBuffer(w * count, np.uint32)
    args = (dc.ptr, da.ptr, db.ptr, n, L, count, stream)
    ms = timed(ev, stream, lambda: capi.call("ya32_poly_multiply_many", *args), reps)
    c0 = dc.download(w)
    mm = modmuls_per_polymul(n, L)
    bits = sum(p.bit_length() for p in primes)
    for d in (da, db, dc):
        d.close()
    return {"degree": n, "nmoduli": L, "modulus_bits": bits, "polys_per_launch": count,
            "ms_per_launch": ms, "value": count / (ms / 1e3), "unit": "poly-mults/s",
            "modmuls_T_per_s": mm * count / (ms / 1e3) / 1e12,
            "algorithmic_GBps": 3 * w * 4 * count / (ms / 1e3) / 1e9,
            "_a0": a[:w], "_b0": b[:w], "_c0": c0}


def u32_cpu_reference(u32, seconds=3.0):
    """The reference's 32-bit build (oracle/_ref/libyellref32.so) on the same
    first poly: one thread, bounded sample, and bit-exactness of the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle  # checker only
    a0, b0, c0 = u32.pop("_a0"), u32.pop("_b0"), u32.pop("_c0")
    n, L = u32["degree"], u32["nmoduli"]
    try:
        ref = _oracle.RefLib(32)
    except RuntimeError:
        orc = _oracle.Oracle(32)
        return {"kind": "port", "bit_exact": bool(np.array_equal(orc.polymul(a0, b0, n, L), c0))}
    want = ref.polymul(a0, b0, n, L)
    lg = n.bit_length() - 1
    t1 = ref.lib.yref_bench_polymul(lg, capi.ptr(a0), capi.ptr(b0), L, 1, 1)
    reps = max(1, int(seconds / max(t1, 1e-6)))
    s1 = ref.lib.yref_bench_polymul(lg, capi.ptr(a0), capi.ptr(b0), L, 1, reps)
    return {"kind": "reference", "cores": 1, "value": reps / s1, "unit": "poly-mults/s",
            "sample": "%d poly-mults (N=%d, L=%d) of the reference's 32-bit build, one thread" % (reps, n, L),
            "bit_exact": bool(np.array_equal(want, c0))}


def latency_bench(stream, ev, n, L, reps=20):
    """Device time of ONE poly-multiplication (a single N = 2^14, L = 8 request) as a split
    launch (each residue over many workgroups) and as the one-workgroup-per-
    residue fused kernel; median over reps."""
    a = capi.DeviceBuffer.from_host(capi.sample_uniform(n, L, 31))
    b = capi.DeviceBuffer.from_host(capi.sample_gaussian(n, L, 3.2, 32))
    c = capi.DeviceBuffer(n * L)
    out = {}
    for name, mode in (("split_auto", 0), ("fused", -1)):
        capi.call("ya_split_policy", mode)
        timed(ev, stream, lambda: capi.call("ya_poly_multiply", c.ptr, a.ptr, b.ptr, n, L, stream), 1)
        ts = []
        for _ in range(reps + 2):
            capi.call("ya_stream_synchronize", stream)
            capi.call("ya_event_record", ev.a, stream)
            capi.call("ya_poly_multiply", c.ptr, a.ptr, b.ptr, n, L, stream)
            capi.call("ya_event_record", ev.b, stream)
            ts.append(ev.elapsed_ms() * 1e3)
        out[name + "_us"] = float(np.median(ts[2:]))
    capi.call("ya_split_policy", 0)
    for d in (a, b, c):
        d.close()
    return out


def crt_bench(stream, ev, reps=10):
    """CRT lift (ya_poly_crt_lift) of one polynomial, device-resident in and out."""
    out = {}
    for n, L in ((1 << 14, 8), (1 << 16, 24)):
        w = ctypes.c_size_t()
        capi.call("ya_crt_limbs", L, ctypes.byref(w))
        res = capi.sample_uniform(n, L, 77 + L)
        din = capi.DeviceBuffer.from_host(res)
        dout = capi.DeviceBuffer(n * w.value)
        ms = timed(ev, stream, lambda: capi.call("ya_poly_crt_lift", dout.ptr, din.ptr, n, L, stream), reps)
        alg = (n * L + n * w.value) * 8
        out["n%d_l%d" % (n, L)] = {"us_per_poly": ms * 1e3, "coeffs_per_s": n / (ms / 1e3),
                                   "limbs": w.value, "algorithmic_bytes": alg,
                                   "hbm_GBps_algorithmic": alg / (ms / 1e3) / 1e9}
        if n == 1 << 14:
            out["_c2_in"], out["_c2_out"] = res, dout.download().reshape(n, w.value)
        din.close()
        dout.close()
    return out


def crt_cpu_reference(crt):
    """The reference's poly2mpz (GMP, one thread) on the same N = 2^14, L = 8 polynomial,
    and bit-exactness of the GPU lift against it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle  # checker only
    res, got = crt.pop("_c2_in"), crt.pop("_c2_out")
    try:
        ref = _oracle.RefGmpLib()
    except RuntimeError:
        return {"available": False}
    n, L = 1 << 14, 8
    want = ref.poly2mpz(res, n, L, got.shape[1])
    sec = ref.bench(res, n, L, 5)
    return {"available": True, "kind": "reference", "ms_per_poly": sec * 1e3, "cores": 1,
            "sample": "5 x poly<16384>(8).poly2mpz() of the reference (lib/gmp.cpp, GMP)",
            "bit_exact": bool(np.array_equal(want, got))}


def sampler_bench(stream, ev, reps=20):
    """Device samplers drawing an N = 2^14, L = 8 operand pair straight into HBM (a uniform,
    b Gaussian sigma = 3.2; N = 2^14, L = 8) against the host samplers plus
    the upload they replace; ZO and hwt at the RLWE flow's shape."""
    n, L = 1 << 14, 8
    da, db = capi.DeviceBuffer(n * L), capi.DeviceBuffer(n * L)
    out = {"degree": n, "nmoduli": L}
    runs = {"uniform": lambda: capi.sample_device("uniform", da, n, L, None, 1, stream=stream),
            "gaussian": lambda: capi.sample_device("gaussian", db, n, L, 3.2, 2, stream=stream)}
    for name, fn in runs.items():
        fn()
        capi.call("ya_stream_synchronize", stream)
        ms = timed(ev, stream, fn, reps)
        out[name + "_us"] = ms * 1e3
        out[name + "_hbm_write_GBps"] = n * L * 8 / (ms / 1e3) / 1e9
    out["pair_us"] = out["uniform_us"] + out["gaussian_us"]
    ha, hb = capi.sample_host("uniform", n, L, None, 1), capi.sample_host("gaussian", n, L, 3.2, 2)
    out["device_equals_host"] = bool(np.array_equal(da.download(), ha) and np.array_equal(db.download(), hb))
    t0 = time.perf_counter()
    hreps = 3
    for _ in range(hreps):
        ha, hb = capi.sample_host("uniform", n, L, None, 1), capi.sample_host("gaussian", n, L, 3.2, 2)
        da.upload(ha, stream)
        db.upload(hb, stream)
    capi.call("ya_stream_synchronize", stream)
    out["host_sample_plus_upload_us"] = (time.perf_counter() - t0) / hreps * 1e6
    n2, L2 = 8192, 4
    dz = capi.DeviceBuffer(n2 * L2)
    for name, kind, param in (("zo_n8192_l4", "zo", 0x7F), ("hwt64_n8192_l4", "hwt", 64)):
        fn = lambda: capi.sample_device(kind, dz, n2, L2, param, 3, stream=stream)
        fn()
        capi.call("ya_stream_synchronize", stream)
        out[name + "_us"] = timed(ev, stream, fn, reps) * 1e3
    for d in (da, db, dz):
        d.close()
    return out


def rlwe_bench(stream, ev, batch=256, reps=5):
    """The reference example's RLWE flow (example.cpp:12-58) at its own
    parameters (N = 8192, L = 4, hwt 64, sigma 3.2): key generation once
    (eager), then a step of `batch` encryptions + decryptions under that key
    (one launch per flow step over all messages: the samplers, the forward of
    u, and the fused steps -- message addition, forward_lazy and
    add_product_of in one pass per ciphertext half, add_product_of and
    backward in one pass for the decryption), captured as one graph and
    replayed; samples from the device samplers.  The eager step is also timed
    with those steps as separate launches, and with pairing off."""
    from yell_amd.rlwe import Rlwe
    n, L = 8192, 4
    f = Rlwe(n, L, stream=stream)
    sx, bx, ax = f.buffer(), f.buffer(), f.buffer()
    f.keygen(sx, bx, ax, (1, 2, 3))  # first use: tables, Gaussian table
    capi.call("ya_stream_synchronize", stream)
    capi.call("ya_event_record", ev.a, stream)
    f.keygen(sx, bx, ax, (1, 2, 3))
    capi.call("ya_event_record", ev.b, stream)
    keygen_us = ev.elapsed_ms() * 1e3
    msgs = f.buffer(batch)
    capi.call("ya_sample_uniform_device_many", msgs.ptr, n, L, batch, 1000, stream)
    cb, ca, u, pl = (f.buffer(batch) for _ in range(4))
    seeds = (10, 10000, 20000)

    def step(fused=True):
        f.encrypt_many(cb, ca, u, msgs, bx, ax, batch, seeds, fused=fused)
        f.decrypt_many(pl, cb, ca, sx, batch, fused=fused)

    step()
    capi.call("ya_stream_synchronize", stream)
    eager_ms = timed(ev, stream, step, 1)
    step(False)  # the steps after each sampler as separate launches, for comparison
    eager_unfused_ms = timed(ev, stream, lambda: step(False), 1)
    capi.call("ya_pair_forward", 0)  # and one residue per workgroup in the transforms
    step(False)
    eager_unpaired_ms = timed(ev, stream, lambda: step(False), 1)
    capi.call("ya_pair_forward", 1)
    step()
    capi.call("ya_graph_begin", stream)
    step()
    g = ctypes.c_void_p()
    capi.call("ya_graph_end", stream, ctypes.byref(g))
    capi.call("ya_graph_launch", g.value, stream)
    capi.call("ya_stream_synchronize", stream)
    ms = timed(ev, stream, lambda: capi.call("ya_graph_launch", g.value, stream), reps)
    capi.call("ya_graph_destroy", g.value)
    w = n * L
    first = {"plain0": pl.download(w), "ct0": (cb.download(w), ca.download(w)),
             "sx": sx.download(), "pk": (bx.download(), ax.download()), "seeds": seeds}
    for b in (sx, bx, ax, msgs, cb, ca, u, pl):
        b.close()
    return {"degree": n, "nmoduli": L, "hwt": 64, "sigma": 3.2, "batch": batch, "keygen_us": keygen_us,
            "eager_step_ms": eager_ms, "eager_step_unfused_ms": eager_unfused_ms,
            "eager_step_unpaired_unfused_ms": eager_unpaired_ms,
            "graph_step_ms": ms, "value": batch / (ms / 1e3),
            "unit": "encrypt+decrypt/s", "launches_per_step": 7, "launches_per_step_unfused": 12, "_first": first}


def rlwe_cpu_reference(rl, seconds=3.0):
    """The same flow through the reference's own sources (oracle/_ref) on the
    samples of the device run (host samplers, same seeds): bit-exactness of
    message 0's key, ciphertext and decryption, then encrypt + decrypt rounds
    timed on one thread and on all threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle  # checker only
    first = rl.pop("_first")
    n, L = rl["degree"], rl["nmoduli"]
    h = lambda kind, param, s: capi.sample_host(kind, n, L, param, s, 64)
    sx, e0, ax = h("hwt", 64, 1), h("gaussian", 3.2, 2), h("uniform", None, 3)
    su, se1, se2 = first["seeds"]
    msg, u, e1, e2 = h("uniform", None, 1000), h("zo", 0x7F, su), h("gaussian", 3.2, se1), h("gaussian", 3.2, se2)
    try:
        ref = _oracle.RefLib(64)
    except RuntimeError:
        orc = _oracle.Oracle(64)
        got = _oracle.rlwe_oracle().flow(orc, n, L, sx, e0, ax, msg, u, e1, e2)
        return {"kind": "port", "bit_exact": bool(np.array_equal(got[4], first["plain0"]))}
    sxf, pk_bx, ct_bx, ct_ax, plain = ref.rlwe(n, L, sx, e0, ax, msg, u, e1, e2)
    exact = all(np.array_equal(x, y) for x, y in ((sxf, first["sx"]), (pk_bx, first["pk"][0]),
                                                   (ax, first["pk"][1]), (ct_bx, first["ct0"][0]),
                                                   (ct_ax, first["ct0"][1]), (plain, first["plain0"])))
    t1 = ref.bench_rlwe(n, L, sxf, pk_bx, ax, msg, u, e1, e2, 1, 2) / 2
    reps = max(1, int(seconds / max(t1, 1e-6)))
    s1 = ref.bench_rlwe(n, L, sxf, pk_bx, ax, msg, u, e1, e2, 1, reps)
    th = min(16, os.cpu_count() or 1)
    reps_t = max(1, int(seconds / max(t1, 1e-6)))
    st = ref.bench_rlwe(n, L, sxf, pk_bx, ax, msg, u, e1, e2, th, reps_t)
    return {"kind": "reference", "bit_exact": bool(exact), "unit": "encrypt+decrypt/s",
            "value_1core": reps / s1, "value": th * reps_t / st, "cores": th,
            "sample": "%d encrypt+decrypt rounds on 1 thread, then %d on each of %d threads, of example.cpp's "
                      "flow (N=%d, L=%d) through the reference's sources" % (reps, reps_t, th, n, L)}


def served_bench(rates, duration, shapes, pool=8, max_pending=1024, batch_jobs=256, modes=("batched", "plain"),
                 inflight=3, verify=1):
    """Poisson stream of host-operand requests (shape drawn uniformly from
    `shapes`, `pool` operand pairs per shape): continuous batching vs one
    request per launch, same seeded stream."""
    out = {"shapes": [list(x) for x in shapes] if len(shapes) <= 4 else "%d shapes" % len(shapes),
           "duration_s": duration, "runs": []}
    degrees = (ctypes.c_uint32 * len(shapes))(*[n for n, _ in shapes])
    mods = (ctypes.c_uint32 * len(shapes))(*[L for _, L in shapes])
    configs = {"batched": dict(max_batch_jobs=batch_jobs, max_inflight=inflight),
               "plain": dict(max_batch_jobs=1, max_inflight=1)}
    for mode, cfgkw in ((m, configs[m]) for m in modes):
        cfg = capi.ServerConfig(hbm_budget_bytes=2 << 30, max_pending=max_pending, max_moduli=16, **cfgkw)
        srv = ctypes.c_void_p()
        capi.call("ya_server_create", ctypes.byref(srv), ctypes.byref(cfg))
        # untimed warm-up stream (first-use costs of the server's streams,
        # arena and pinned pools stay out of the measured percentiles)
        warm = capi.PoissonSpec(rate_per_s=rates[0], duration_s=0.3, nshapes=len(shapes), degrees=degrees,
                                nmoduli=mods, pool=pool, verify=0, seed=7)
        capi.call("ya_server_run_poisson", srv, ctypes.byref(warm), ctypes.byref(capi.ServeReport()))
        for rate in rates:
            spec = capi.PoissonSpec(rate_per_s=rate, duration_s=duration, nshapes=len(shapes),
                                    degrees=degrees, nmoduli=mods, pool=pool, verify=verify, seed=99)
            rep = capi.ServeReport()
            log("served", mode, "rate", rate)
            capi.call("ya_server_run_poisson", srv, ctypes.byref(spec), ctypes.byref(rep))
            d = rep.as_dict()
            d.update({"mode": mode, "offered_rate_per_s": rate})
            out["runs"].append(d)
        capi.call("ya_server_destroy", srv)
    return out


KNEE_P99_US = 5000.0


def served_sweep(rates, duration, n, L, batch_jobs=256, inflight=3, verify=1):
    """Batched served stream (N = 2^14, L = 8, every result verified) at increasing
    offered loads until it saturates: the knee is the highest offered rate
    whose completed throughput keeps within 3% of it with no rejection, no
    failure and p99 latency within KNEE_P99_US."""
    out = served_bench(rates, duration, [(n, L)], modes=("batched",), batch_jobs=batch_jobs, inflight=inflight,
                       verify=verify)
    out["knee_rule"] = "throughput >= 0.97 x offered, no rejection or failure, p99 <= %.0f us" % KNEE_P99_US
    knee = None
    for r in out["runs"]:
        # against the arrivals the Poisson draw actually made in the window
        ok = (r["rejected"] == 0 and r["failed"] == 0 and r["mismatches"] == 0 and
              r["throughput_per_s"] >= 0.97 * r["offered"] / duration and r["p99_us"] <= KNEE_P99_US)
        r["sustained"] = bool(ok)
        if ok and (knee is None or r["offered_rate_per_s"] > knee):
            knee = r["offered_rate_per_s"]
    out["knee_rate_per_s"] = knee
    out["max_throughput_per_s"] = max(r["throughput_per_s"] for r in out["runs"])
    at = [r for r in out["runs"] if r["offered_rate_per_s"] == knee]
    if at:
        out["p50_us_at_knee"], out["p99_us_at_knee"] = at[0]["p50_us"], at[0]["p99_us"]
        out["jobs_per_batch_at_knee"] = at[0]["mean_batch_jobs"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the uneven batch, probes and tuner")
    ap.add_argument("--only", default="",
                    help="comma list of the extra sections to run (tune, probe, c3, c4, crt, latency, u32, "
                         "samplers, rlwe, sweep, served); default all")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[4000.0, 8000.0, 12000.0],
                    help="offered loads (requests/s) of the served-form runs")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--sweep-rates", type=float, nargs="+",
                    default=[4000.0, 8000.0, 12000.0, 16000.0, 20000.0, 24000.0, 28000.0],
                    help="offered loads of the batched served sweep to the saturation knee")
    ap.add_argument("--sweep-seconds", type=float, default=1.0)
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X and this path does not "
              "shard across GPUs (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    n, L, B = 1 << 14, 8, args.batch
    w = n * L
    capi.load()
    capi.call("ya_init", 0)
    stream = ctypes.c_void_p()
    capi.call("ya_stream_create", ctypes.byref(stream))
    ev = Events()
    extra = {}
    if args.copy_calib_mb:
        nw = args.copy_calib_mb * (1 << 20) // 8
        s_, d_ = capi.DeviceBuffer(nw), capi.DeviceBuffer(nw)
        capi.call("ya_memcpy_dtod", d_.ptr, s_.ptr, nw * 8, stream)
        capi.call("ya_stream_synchronize", stream)
        s_.close()
        d_.close()
    only = set(x for x in args.only.split(",") if x)

    def want(section):
        return not args.no_extra and (not only or section in only)

    if want("tune"):
        log("autotune 2^10..2^15")
        S = 7  # YA_TUNE_REPORT_STRIDE
        rep = (ctypes.c_double * (6 * S))()
        capi.call("ya_autotune", 10, 15, rep)
        lib_ = capi.load()
        extra["autotune"] = {"classes": list(range(10, 16)),
                             "ms_batched_r8_r16_fused_r8_r16_r32_fused_two_sequences_r8_r16": [
                                 [round(rep[S * k + j], 4) for j in range(S)] for k in range(6)],
                             "chosen_batched_le": [lib_.ya_tune_get(2, lg) for lg in range(10, 16)],
                             "chosen_fused_le": [lib_.ya_tune_get(1, lg) for lg in range(10, 16)],
                             "chosen_fused_dual": [lib_.ya_tune_get(3, lg) for lg in range(10, 16)]}
    if want("probe"):
        probes = {}
        for kind, name in ((0, "mad_u64_u32_per_s"), (1, "mul_lo_u32_per_s"), (2, "ct_butterflies_per_s"),
                           (3, "ct32_butterflies_per_s")):
            v = ctypes.c_double()
            capi.call("ya_probe_valu", kind, ctypes.byref(v))
            probes[name] = v.value
        extra["valu_probe"] = probes

    log("inputs: batch", B, "N", n, "L", L)
    a, b = make_inputs(n, L, B, seed=1234)
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    dc = capi.DeviceBuffer(B * w)
    capi.call("ya_ntt_init_tables", n, L)
    plan = make_plan([(da.ptr + k * w * 8, db.ptr + k * w * 8, dc.ptr + k * w * 8, n, L)
                      for k in range(B)])
    form = ctypes.c_int()
    capi.call("ya_batch_plan_info", plan, None, None, None, ctypes.byref(form))
    strided = ctypes.c_int()
    capi.call("ya_batch_plan_is_strided", plan, ctypes.byref(strided))
    # a strided plan runs the looping kernel when the batch gives every CU's
    # one 2^14 workgroup at least two residues
    looping = B * L >= 2 * CUS
    step_kernel = (("k_polymul_loop<14,4>" if looping else "k_polymul<14,4,dual>") if strided.value
                   else ("k_polymul_tab_loop<14,4>" if looping else "k_polymul_tab<14,4,dual>")) \
        if form.value == 2 else "k_batch_polymul"

    def batched():
        capi.call("ya_batch_plan_run_graph", plan, stream)

    def grid_form():
        capi.call("ya_poly_multiply_many", dc.ptr, da.ptr, db.ptr, n, L, B, stream)

    def plain():
        for k in range(B):
            off = k * w * 8
            capi.call("ya_poly_multiply", dc.ptr + off, da.ptr + off, db.ptr + off, n, L, stream)

    def settle_on(fn, seconds=0.2):
        # after host-side work (input generation, a large download) the GPU
        # has idled and its clocks ramp back up over ~0.1 s of work; keep
        # stepping (untimed) until `seconds` of steps have run
        k = 0
        t_settle = time.perf_counter()
        while time.perf_counter() - t_settle < seconds:
            fn()
            capi.call("ya_stream_synchronize", stream)
            k += 1
        return k

    for _ in range(args.warmup):
        batched()
    capi.call("ya_stream_synchronize", stream)
    settle = settle_on(batched)
    log("timing batched:", args.steps, "steps")
    t0 = time.perf_counter()
    ms_step = timed(ev, stream, batched, args.steps)
    wall = time.perf_counter() - t0
    value = B / (ms_step / 1e3)
    # every output word of the three forms is compared (the buffer is cleared
    # between them, so each form writes all of it)
    c_batched = dc.download()
    capi.call("ya_memset_zero", dc.ptr, B * w * 8, stream)
    grid_form()
    capi.call("ya_stream_synchronize", stream)
    same_grid = bool(np.array_equal(c_batched, dc.download()))
    settle_on(grid_form, 0.1)
    ms_grid = timed(ev, stream, grid_form, max(1, args.steps // 2))
    capi.call("ya_memset_zero", dc.ptr, B * w * 8, stream)
    plain()
    capi.call("ya_stream_synchronize", stream)
    log("timing plain:", args.plain_steps, "steps of", B, "launches")
    ms_plain = timed(ev, stream, plain, args.plain_steps)
    plain_value = B / (ms_plain / 1e3)
    same_plain = bool(np.array_equal(c_batched, dc.download()))

    bytes_per = 3 * n * L * 8
    achieved = bytes_per * B / (ms_step / 1e3) / 1e9
    mm = modmuls_per_polymul(n, L)
    out = {
        "metric": METRIC, "value": value, "unit": "poly-mults/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "settle_steps": settle, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic (seeded uniform x Gaussian sigma=3.2 operands)",
        "config": {"workload": "RNS poly-multiplication N=16384 L=8 (P[0..7]), a uniform, "
                               "b Gaussian sigma=3.2; %d polys per batched launch "
                               "(batch plan replayed as a hipGraph)" % B,
                   "launch_form": {1: "persistent", 2: "grid"}.get(form.value, "?"),
                   "degree": n, "nmoduli": L, "batch": B},
        "plain": {"value": plain_value, "unit": "poly-mults/s", "ms_per_step": ms_plain,
                  "launches_per_step": B},
        "batched_over_plain": value / plain_value,
        "grid_form": {"value": B / (ms_grid / 1e3), "ms_per_step": ms_grid,
                      "kernel": ("k_polymul_loop<14,4>, one workgroup per CU" if B * L >= 2 * CUS
                                 else "k_polymul<14,4>, one workgroup per residue")},
        "roofline": roofline(step_kernel, ms_step, B, achieved, bytes_per),
        "modmul_rate": {"modmuls_per_polymul": mm,
                        "achieved_T_per_s": mm * B / (ms_step / 1e3) / 1e12},
        "timed_wall_s": wall,
    }
    out.update(extra)
    if "valu_probe" in extra:
        pk = extra["valu_probe"]["ct_butterflies_per_s"]
        out["modmul_rate"]["butterfly_probe_T_per_s"] = pk / 1e12
        out["modmul_rate"]["frac_of_probe"] = mm * B / (ms_step / 1e3) / pk
    if want("c3"):
        log("uneven batch: 1024 jobs")
        out["c3_uneven_batch"] = c3_bench(stream, ev, max(2, args.steps // 4))
    if want("c4"):
        log("large-ring MAC: N = 2^16, L = 24, 8 pairs")
        out["c4_large_ring_mac"] = c4_bench(stream, ev, 3)
    if want("crt"):
        log("CRT lift")
        out["crt_lift"] = crt_bench(stream, ev)
    if want("latency"):
        log("single poly-mult latency")
        out["single_polymul_latency"] = latency_bench(stream, ev, n, L)
    if want("u32"):
        log("32-bit moduli build")
        out["u32_moduli"] = u32_bench(stream, ev)
        if "valu_probe" in extra:  # the 32-bit build against its own butterfly ceiling
            pk32 = extra["valu_probe"]["ct32_butterflies_per_s"]
            out["u32_moduli"]["butterfly_probe_T_per_s"] = pk32 / 1e12
            out["u32_moduli"]["frac_of_probe"] = out["u32_moduli"]["modmuls_T_per_s"] * 1e12 / pk32
    if want("samplers"):
        log("device samplers")
        out["device_samplers"] = sampler_bench(stream, ev)
    if want("rlwe"):
        log("RLWE flow")
        out["rlwe_flow"] = rlwe_bench(stream, ev)
    if want("sweep"):
        log("served sweep")
        sw = served_sweep(args.sweep_rates, args.sweep_seconds, n, L)
        out["served_sweep"] = sw
        out["served_knee_rate_per_s"] = sw["knee_rate_per_s"]
        out["served_p99_us_at_knee"] = sw.get("p99_us_at_knee")
    if want("served"):
        sv = served_bench(args.serve_rates, args.serve_seconds, [(n, L)])
        out["served"] = sv
        for r in sv["runs"]:
            if r["mode"] == "batched" and r["offered_rate_per_s"] == args.serve_rates[0]:
                out["served_p50_us"], out["served_p99_us"] = r["p50_us"], r["p99_us"]
                out["served_offered_rate_per_s"] = r["offered_rate_per_s"]
        # the same stream over the uneven batch's shapes (degree 2^10..2^15
        # and 1..12 moduli, each uniform): 72 shapes, 2 operand pairs each
        uneven = [(1 << lg, m) for lg in range(10, 16) for m in range(1, 13)]
        out["served_uneven"] = served_bench(sorted({args.serve_rates[0], args.serve_rates[-1]}), args.serve_seconds,
                                            uneven, pool=2, max_pending=256, batch_jobs=64)
    same = same_grid and same_plain
    out["batched_equals_plain"] = same
    out["compared"] = {"batched_vs_grid_and_plain_words": B * w, "batched_equals_grid": same_grid,
                       "batched_equals_plain": same_plain}
    if not args.no_cpu:
        log("cpu baseline (~%.0f s)" % args.cpu_seconds)
        sample = sorted({0, B - 1} | {int(k) for k in np.linspace(0, B - 1, 8)})
        cb, exact = cpu_baseline(a, b, n, L, c_batched, sample, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["compared"]["polys_vs_cpu_reference"] = sample
        out["bit_exact_vs_cpu"] = exact and same
        out["gpu_over_cpu"] = value / cb["value"]
        if "crt_lift" in out:
            log("CRT lift cpu reference")
            out["crt_lift"]["cpu_reference"] = crt_cpu_reference(out["crt_lift"])
        if "u32_moduli" in out:
            log("32-bit cpu reference")
            out["u32_moduli"]["cpu_reference"] = u32_cpu_reference(out["u32_moduli"])
        if "rlwe_flow" in out:
            log("RLWE cpu reference")
            out["rlwe_flow"]["cpu_reference"] = rlwe_cpu_reference(out["rlwe_flow"])
    if "crt_lift" in out:
        out["crt_lift"].pop("_c2_in", None)
        out["crt_lift"].pop("_c2_out", None)
    for k in ("_a0", "_b0", "_c0"):
        out.get("u32_moduli", {}).pop(k, None)
    out.get("rlwe_flow", {}).pop("_first", None)
    capi.load().ya_batch_plan_destroy(plan)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
