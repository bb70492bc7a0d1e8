"""Batched form: uneven jobs in one persistent launch (and as a hipGraph),
bit-exact against the oracle and against the plain per-poly path."""
import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from _oracle import Oracle
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
orc = Oracle()


def _jobs(rng, njobs, logs, lmax):
    specs = [(1 << int(rng.choice(logs)), int(rng.integers(1, lmax + 1))) for _ in range(njobs)]
    host = []
    for k, (n, L) in enumerate(specs):
        host.append((n, L, capi.sample_uniform(n, L, 1000 + 2 * k),
                     capi.sample_uniform(n, L, 1001 + 2 * k)))
    return host


class Batch:
    def __init__(self, host):
        device()
        self.host = host
        self.bufs = []
        arr = (capi.PolymulJob * len(host))()
        for k, (n, L, a, b) in enumerate(host):
            da, db, dc = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b), capi.DeviceBuffer(n * L)
            self.bufs.append((da, db, dc))
            arr[k] = capi.PolymulJob(da.ptr, db.ptr, dc.ptr, n, L)
        self.jobs = arr
        self.plan = ctypes.c_void_p()
        capi.call("ya_batch_plan_create", ctypes.byref(self.plan), ctypes.cast(arr, ctypes.c_void_p), len(host))

    def results(self):
        return [dc.download() for _, _, dc in self.bufs]

    def close(self):
        capi.load().ya_batch_plan_destroy(self.plan)


def test_uneven_batch_vs_oracle():
    rng = np.random.default_rng(1)
    host = _jobs(rng, 30, [7, 10, 11, 12, 13, 14, 15], 12)
    say("batch uneven", [(n, L) for n, L, _, _ in host])
    bt = Batch(host)
    items, res, wgs = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    capi.call("ya_batch_plan_info", bt.plan, ctypes.byref(items), ctypes.byref(res), ctypes.byref(wgs), None)
    assert res.value == sum(L for _, L, _, _ in host)
    for rep in range(2):   # the queue resets itself between launches
        capi.call("ya_batch_plan_run", bt.plan, None)
        capi.call("ya_device_synchronize")
        for (n, L, a, b), got in zip(host, bt.results()):
            np.testing.assert_array_equal(got, orc.polymul(a, b, n, L))
    bt.close()


def test_graph_replays_with_new_data():
    rng = np.random.default_rng(2)
    host = _jobs(rng, 12, [10, 12, 14, 15], 6)
    bt = Batch(host)
    for rep in range(3):
        say("graph replay", rep)
        new = []
        for k, ((n, L, _, _), (da, db, _)) in enumerate(zip(host, bt.bufs)):
            a = capi.sample_uniform(n, L, 7000 + 10 * rep + k)
            b = capi.sample_gaussian(n, L, 3.2, 8000 + 10 * rep + k)
            da.upload(a)
            db.upload(b)
            new.append((n, L, a, b))
        capi.call("ya_batch_plan_run_graph", bt.plan, None)
        capi.call("ya_device_synchronize")
        for (n, L, a, b), got in zip(new, bt.results()):
            np.testing.assert_array_equal(got, orc.polymul(a, b, n, L))
    bt.close()


def test_full_c3_batch_vs_oracle_through_graph():
    """The uneven batch's shape: 1024 jobs, degrees 2^10..2^15, 1..12 moduli,
    uniform operands, packed back to back in three buffers, run as the plan's
    hipGraph (captured, then replayed with new operands): every job of both
    runs against the C oracle, and the first run against the plain
    ya_poly_multiply path."""
    device()
    rng = np.random.default_rng(2026)
    njobs = 1024
    specs = [(1 << int(rng.integers(10, 16)), int(rng.integers(1, 13))) for _ in range(njobs)]
    offs = np.cumsum([0] + [n * L for n, L in specs])
    total = int(offs[-1])
    say("uneven 1024-job batch through the graph", total, "words per operand")
    A, B, C, D = (capi.DeviceBuffer(total) for _ in range(4))
    arr = (capi.PolymulJob * njobs)()
    for k, (n, L) in enumerate(specs):
        o = 8 * int(offs[k])
        arr[k] = capi.PolymulJob(A.ptr + o, B.ptr + o, C.ptr + o, n, L)
    plan = ctypes.c_void_p()
    capi.call("ya_batch_plan_create", ctypes.byref(plan), ctypes.cast(arr, ctypes.c_void_p), njobs)
    stream = ctypes.c_void_p()
    capi.call("ya_stream_create", ctypes.byref(stream))

    def operands(seed):
        a = np.concatenate([capi.sample_uniform(n, L, seed + 2 * k) for k, (n, L) in enumerate(specs)])
        b = np.concatenate([capi.sample_uniform(n, L, seed + 2 * k + 1) for k, (n, L) in enumerate(specs)])
        return a, b

    def check(a, b, got, tag):
        def one(k):
            n, L = specs[k]
            s = slice(int(offs[k]), int(offs[k + 1]))
            return k if not np.array_equal(got[s], orc.polymul(a[s], b[s], n, L)) else -1
        with ThreadPoolExecutor(max_workers=16) as ex:
            bad = [k for k in ex.map(one, range(njobs)) if k >= 0]
        assert not bad, "%s: jobs %s differ from the oracle" % (tag, bad[:10])

    try:
        for run, seed in enumerate((50000, 70000)):
            a, b = operands(seed)
            A.upload(a)
            B.upload(b)
            say("uneven batch graph run", run)
            capi.call("ya_batch_plan_run_graph", plan, stream)
            capi.call("ya_stream_synchronize", stream)
            got = C.download()
            check(a, b, got, "graph run %d" % run)
            if run == 0:
                for k, (n, L) in enumerate(specs):
                    o = 8 * int(offs[k])
                    capi.call("ya_poly_multiply", D.ptr + o, A.ptr + o, B.ptr + o, n, L, stream)
                capi.call("ya_stream_synchronize", stream)
                np.testing.assert_array_equal(D.download(), got)
    finally:
        capi.load().ya_batch_plan_destroy(plan)
        capi.call("ya_stream_destroy", stream)
        for d in (A, B, C, D):
            d.close()


def test_rejects_unsupported_degree():
    device()
    n = 1 << 16
    d = capi.DeviceBuffer(3 * n)
    arr = (capi.PolymulJob * 1)(capi.PolymulJob(d.ptr, d.ptr + 8 * n, d.ptr + 16 * n, n, 1))
    plan = ctypes.c_void_p()
    rc = capi.load().ya_batch_plan_create(ctypes.byref(plan), ctypes.cast(arr, ctypes.c_void_p), 1)
    assert rc == 5   # YA_ERR_UNSUPPORTED
    arr[0] = capi.PolymulJob(d.ptr, d.ptr + 8 * 1024, d.ptr, 1024, 1)   # c aliases a
    assert capi.load().ya_batch_plan_create(ctypes.byref(plan), ctypes.cast(arr, ctypes.c_void_p), 1) == 1


def test_every_pass_shape_gives_the_same_words():
    """The autotuner may pick radix-8 or radix-16 passes: both give the oracle's
    words, in every sub-group form of the persistent kernel (one wave or less:
    no barrier between passes; larger sub-groups: the workgroup's barrier)."""
    lib = device()
    rng = np.random.default_rng(5)
    host = _jobs(rng, 14, [7, 8, 9, 10, 11, 12, 13], 5)
    saved = {lg: lib.ya_tune_get(2, lg) for lg in range(7, 14)}
    try:
        for le in (3, 4):
            for lg in range(7, 14):
                capi.call("ya_tune_set", 2, lg, le)
            say("pass shape le", le)
            bt = Batch(host)
            capi.call("ya_batch_plan_run", bt.plan, None)
            capi.call("ya_device_synchronize")
            for (n, L, a, b), got in zip(host, bt.results()):
                np.testing.assert_array_equal(got, orc.polymul(a, b, n, L))
            bt.close()
    finally:
        for lg, le in saved.items():
            capi.call("ya_tune_set", 2, lg, le)


def test_staggered_start_gives_the_same_words():
    """The first wave's staggered start (tune kind 5) only delays workgroups:
    at 0, 3 and 8 steps the fused grid kernel (2,560 residues of 2^14, ten
    waves) and the persistent kernel (2,560 items of 2^14, ten per
    workgroup) give the same words, and sampled polys the oracle's."""
    lib = device()
    n, L, count = 1 << 14, 8, 320
    a = np.concatenate([capi.sample_uniform(n, L, 7100 + k) for k in range(count)])
    b = np.concatenate([capi.sample_uniform(n, L, 7300 + k) for k in range(count)])
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    dc = capi.DeviceBuffer(n * L * count)
    rng = np.random.default_rng(13)
    host = [(n, L, capi.sample_uniform(n, L, 3000 + 2 * k), capi.sample_uniform(n, L, 3001 + 2 * k))
            for k in range(count)]
    saved = {lg: lib.ya_tune_get(5, lg) for lg in (14, 15)}
    capi.call("ya_batch_force_form", 1)
    try:
        bt = Batch(host)
        got_grid, got_pers = {}, {}
        for steps in (0, 3, 8):
            for lg in (14, 15):
                capi.call("ya_tune_set", 5, lg, steps)
            assert lib.ya_tune_get(5, 14) == steps
            say("stagger steps", steps)
            capi.call("ya_poly_multiply_many", dc.ptr, da.ptr, db.ptr, n, L, count, None)
            capi.call("ya_batch_plan_run", bt.plan, None)
            capi.call("ya_device_synchronize")
            got_grid[steps] = dc.download()
            got_pers[steps] = np.concatenate(bt.results())
        for steps in (3, 8):
            np.testing.assert_array_equal(got_grid[steps], got_grid[0])
            np.testing.assert_array_equal(got_pers[steps], got_pers[0])
        w = n * L
        for k in (0, int(rng.integers(1, count - 1)), count - 1):
            np.testing.assert_array_equal(got_grid[0][k * w:(k + 1) * w],
                                          orc.polymul(a[k * w:(k + 1) * w], b[k * w:(k + 1) * w], n, L))
        for k in (0, count - 1):
            _, _, ha, hb = host[k]
            np.testing.assert_array_equal(got_pers[0][k * w:(k + 1) * w], orc.polymul(ha, hb, n, L))
        bt.close()
    finally:
        capi.call("ya_batch_force_form", 0)
        for lg, v in saved.items():
            capi.call("ya_tune_set", 5, lg, v)
    assert lib.ya_tune_set(5, 14, 17) != 0 and lib.ya_tune_set(5, 14, -1) != 0


def test_autotune_runs_and_fills_table():
    """Every timed shape of 2^12 and 2^13 reports a time: batched radix 8 / 16,
    fused radix 8 / 16 / 32 and the fused two-sequence form at radix 8 / 16;
    the table then holds shapes that exist."""
    lib = device()
    S = 7  # YA_TUNE_REPORT_STRIDE
    rep = (ctypes.c_double * (2 * S))()
    capi.call("ya_autotune", 12, 13, rep)
    assert all(rep[S * c + k] > 0 for c in range(2) for k in range(S))
    assert lib.ya_tune_get(2, 12) in (3, 4) and lib.ya_tune_get(1, 13) in (3, 4, 5)
    assert lib.ya_tune_get(3, 12) in (0, 1) and lib.ya_tune_get(3, 13) in (0, 1)
    for lg in (12, 13):  # leave the defaults for the tests that follow
        for kind in (0, 1, 2):
            capi.call("ya_tune_set", kind, lg, 4)
        capi.call("ya_tune_set", 3, lg, 1)


@pytest.mark.parametrize("form", [1, 2])
def test_both_launch_forms_of_a_uniform_batch(form):
    """A one-class batch as the persistent queue and as the residue-table grid."""
    lib = device()
    rng = np.random.default_rng(11)
    host = _jobs(rng, 40, [14], 8)
    capi.call("ya_batch_force_form", form)
    try:
        bt = Batch(host)
        got_form = ctypes.c_int()
        capi.call("ya_batch_plan_info", bt.plan, None, None, None, ctypes.byref(got_form))
        assert got_form.value == form
        say("launch form", form)
        capi.call("ya_batch_plan_run_graph", bt.plan, None)
        capi.call("ya_device_synchronize")
        for (n, L, a, b), got in zip(host, bt.results()):
            np.testing.assert_array_equal(got, orc.polymul(a, b, n, L))
        bt.close()
    finally:
        capi.call("ya_batch_force_form", 0)


@pytest.mark.parametrize("logn,L,count", [(14, 8, 128), (14, 8, 320), (12, 3, 200), (10, 1, 600)])
def test_contiguous_plan_strided_and_table(logn, L, count):
    """Back-to-back jobs of one shape: the direct-addressed path and the
    residue-table path give the same words, and both the oracle's (first and
    last job), over two graph replays (2^14 x 8 x 320: ten waves of
    workgroups, so the captured launches start staggered)."""
    device()
    n = 1 << logn
    w = n * L
    say("contiguous plan", logn, L, count)
    a = np.concatenate([capi.sample_uniform(n, L, 10 + k) for k in range(count)])
    b = np.concatenate([capi.sample_gaussian(n, L, 3.2, 5000 + k) for k in range(count)])
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    outs = []
    for strided in (1, 0):
        capi.call("ya_batch_strided", strided)
        dc = capi.DeviceBuffer(w * count)
        arr = (capi.PolymulJob * count)(*[capi.PolymulJob(da.ptr + 8 * k * w, db.ptr + 8 * k * w,
                                                          dc.ptr + 8 * k * w, n, L) for k in range(count)])
        plan = ctypes.c_void_p()
        capi.call("ya_batch_plan_create", ctypes.byref(plan), ctypes.cast(arr, ctypes.c_void_p), count)
        form = ctypes.c_int()
        capi.call("ya_batch_plan_info", plan, None, None, None, ctypes.byref(form))
        for _ in range(2):
            capi.call("ya_batch_plan_run_graph", plan, None)
        capi.call("ya_device_synchronize")
        outs.append((form.value, dc.download()))
        capi.load().ya_batch_plan_destroy(plan)
        dc.close()
    capi.call("ya_batch_strided", 1)
    assert outs[0][0] == outs[1][0]
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    for k in (0, count - 1):
        s = slice(k * w, (k + 1) * w)
        np.testing.assert_array_equal(outs[0][1][s], orc.polymul(a[s], b[s], n, L))
    da.close()
    db.close()
