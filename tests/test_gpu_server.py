"""Served form: continuous batching with HBM-budget admission.  Every result
equals the oracle's; a request gives the same words alone and in a mix."""
import ctypes

import numpy as np
import pytest

from _oracle import Oracle
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
orc = Oracle()


class Pinned:
    def __init__(self, nwords):
        p = ctypes.c_void_p()
        capi.call("ya_host_alloc_pinned", ctypes.byref(p), nwords * 8)
        self.p = p.value
        self.a = np.ctypeslib.as_array((ctypes.c_uint64 * nwords).from_address(self.p))

    def close(self):
        capi.load().ya_host_free_pinned(self.p)


def _server(**kw):
    device()
    cfg = capi.ServerConfig(**kw)
    s = ctypes.c_void_p()
    capi.call("ya_server_create", ctypes.byref(s), ctypes.byref(cfg))
    return s


def _submit(s, a, b, c, n, L):
    t = ctypes.c_uint64()
    capi.call("ya_server_submit", s, a.p, b.p, c.p, n, L, ctypes.byref(t))
    return t.value


def _request(rng, k):
    n, L = 1 << int(rng.integers(10, 16)), int(rng.integers(1, 13))
    a, b, c = Pinned(n * L), Pinned(n * L), Pinned(n * L)
    a.a[:] = capi.sample_uniform(n, L, 300 + 2 * k)
    b.a[:] = capi.sample_gaussian(n, L, 3.2, 301 + 2 * k)
    return n, L, a, b, c


def test_requests_match_oracle_and_alone_equals_mix():
    s = _server(hbm_budget_bytes=256 << 20, max_batch_jobs=32, max_inflight=2)
    rng = np.random.default_rng(1)
    # the probe request, alone
    n, L, a, b, c = _request(rng, 0)
    say("server alone", n, L)
    t = _submit(s, a, b, c, n, L)
    capi.call("ya_server_wait", s, t, 60000, None)
    alone = c.a.copy()
    np.testing.assert_array_equal(alone, orc.polymul(a.a, b.a, n, L))
    # the same request inside a mix of 60 others
    others = [_request(rng, k) for k in range(1, 61)]
    c2 = Pinned(n * L)
    tickets = []
    for k, (n2, L2, a2, b2, cc) in enumerate(others):
        tickets.append(_submit(s, a2, b2, cc, n2, L2))
        if k == 30:
            tp = _submit(s, a, b, c2, n, L)
    say("server mix", len(others))
    for t in tickets:
        capi.call("ya_server_wait", s, t, 60000, None)
    capi.call("ya_server_wait", s, tp, 60000, None)
    np.testing.assert_array_equal(c2.a, alone)
    for n2, L2, a2, b2, cc in others:
        np.testing.assert_array_equal(cc.a, orc.polymul(a2.a, b2.a, n2, L2))
    st = capi.ServerStats()
    capi.call("ya_server_get_stats", s, ctypes.byref(st))
    assert st.completed == 62 and st.batches >= 2 and st.p99_us >= st.p50_us > 0
    capi.call("ya_server_destroy", s)


def test_admission_bounds():
    s = _server(hbm_budget_bytes=1 << 20, max_pending=2)
    lib = capi.load()
    n, L = 1 << 14, 8                     # 3 MiB of operands > 1 MiB budget
    big = Pinned(n * L)
    t = ctypes.c_uint64()
    assert lib.ya_server_submit(s, big.p, big.p, big.p, n, L, ctypes.byref(t)) == 6   # REJECTED
    assert lib.ya_server_submit(s, big.p, big.p, big.p, 1 << 16, 1, ctypes.byref(t)) == 5  # 2^16
    st = capi.ServerStats()
    capi.call("ya_server_get_stats", s, ctypes.byref(st))
    assert st.rejected == 1
    capi.call("ya_server_destroy", s)


def test_poisson_stream_verified():
    s = _server(hbm_budget_bytes=512 << 20, max_batch_jobs=64, max_inflight=2)
    degrees = (ctypes.c_uint32 * 2)(1 << 12, 1 << 14)
    mods = (ctypes.c_uint32 * 2)(4, 8)
    spec = capi.PoissonSpec(rate_per_s=3000.0, duration_s=0.5, nshapes=2, degrees=degrees,
                            nmoduli=mods, pool=4, verify=1, seed=7)
    rep = capi.ServeReport()
    say("poisson")
    capi.call("ya_server_run_poisson", s, ctypes.byref(spec), ctypes.byref(rep))
    d = rep.as_dict()
    print(d)
    assert d["offered"] > 500 and d["failed"] == 0 and d["mismatches"] == 0
    assert d["completed"] + d["rejected"] == d["offered"]
    assert d["p99_us"] >= d["p50_us"] > 0
    capi.call("ya_server_destroy", s)


class Pageable:
    """An ordinary numpy buffer: not in the library's pinned registry, so the
    server moves it with DMA copies instead of its transfer kernels."""

    def __init__(self, nwords):
        self.a = np.zeros(nwords + 1, dtype=np.uint64)[1:]   # 8-byte aligned, usually not 16
        self.p = self.a.ctypes.data


def test_pinned_and_pageable_requests_in_one_batch():
    """A batch mixing requests in registered pinned memory (moved by the
    transfer kernels), in ordinary pageable memory (DMA copies), and with a
    pinned buffer offset by 8 bytes (not 16-byte aligned: DMA): every result
    equals the oracle's."""
    s = _server(hbm_budget_bytes=256 << 20, max_batch_jobs=32, max_inflight=1)
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(24):
        n, L = 1 << int(rng.integers(10, 15)), int(rng.integers(1, 9))
        kind = k % 3
        if kind == 0:
            a, b, c = Pinned(n * L), Pinned(n * L), Pinned(n * L)
        elif kind == 1:
            a, b, c = Pageable(n * L), Pageable(n * L), Pageable(n * L)
        else:
            a, b, c = Pinned(n * L), Pinned(n * L), Pinned(n * L + 1)
            c.p += 8
            c.a = np.ctypeslib.as_array((ctypes.c_uint64 * (n * L)).from_address(c.p))
        a.a[:] = capi.sample_uniform(n, L, 900 + 2 * k)
        b.a[:] = capi.sample_gaussian(n, L, 3.2, 901 + 2 * k)
        reqs.append((n, L, a, b, c))
    say("server pinned + pageable mix", len(reqs))
    tickets = [_submit(s, a, b, c, n, L) for n, L, a, b, c in reqs]
    for t in tickets:
        capi.call("ya_server_wait", s, t, 60000, None)
    for n, L, a, b, c in reqs:
        np.testing.assert_array_equal(c.a, orc.polymul(a.a, b.a, n, L))
    capi.call("ya_server_destroy", s)
