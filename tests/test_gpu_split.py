"""Split launches: poly-multiplications of few residues spread over 2^k
workgroups per residue (column kernels + block kernels).  Every shape must
give the words of the one-workgroup-per-residue kernel and of the oracle, and
leave the operands untouched."""
import ctypes

import numpy as np
import pytest

from _oracle import Oracle
from gpu_util import device, say
from test_gpu_batch import Batch, _jobs
from yell_amd import capi

pytestmark = pytest.mark.gpu
orc = Oracle()

SHAPES = [(logn, k) for logn in range(11, 15) for k in range(2, 6) if logn - k >= 9]


@pytest.fixture(autouse=True)
def _policy():
    device()
    yield
    capi.call("ya_split_policy", 0)
    capi.call("ya_batch_force_form", 0)


def _multiply(a, b, n, L, count, mode):
    capi.call("ya_split_policy", mode)
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    dc = capi.DeviceBuffer(n * L * count)
    capi.call("ya_poly_multiply_many", dc.ptr, da.ptr, db.ptr, n, L, count, None)
    out, a2, b2 = dc.download(), da.download(), db.download()
    for d in (da, db, dc):
        d.close()
    return out, a2, b2


@pytest.mark.parametrize("logn,k", SHAPES)
def test_split_shape_matches_fused_and_oracle(logn, k):
    n, L, count = 1 << logn, 3, 2
    say("split shape", logn, k)
    a = np.concatenate([capi.sample_uniform(n, L, 10 * logn + j) for j in range(count)])
    b = np.concatenate([capi.sample_gaussian(n, L, 3.2, 20 * logn + j) for j in range(count)])
    got, a2, b2 = _multiply(a, b, n, L, count, k)
    ref, _, _ = _multiply(a, b, n, L, count, -1)
    assert np.array_equal(a2, a) and np.array_equal(b2, b)
    np.testing.assert_array_equal(got, ref)
    w = n * L
    np.testing.assert_array_equal(got[:w], orc.polymul(a[:w], b[:w], n, L))


def test_split_extreme_words():
    """Operands at p - 1 and 0 in every word (the largest lazy values)."""
    n, L = 1 << 14, 8
    ps = [orc.prime(j) for j in range(L)]
    a = np.concatenate([np.full(n, p - 1, dtype=np.uint64) for p in ps])
    b = a.copy()
    b[::3] = 0
    for k in (2, 5):
        got, _, _ = _multiply(a, b, n, L, 1, k)
        np.testing.assert_array_equal(got, orc.polymul(a, b, n, L))


def test_auto_policy_small_and_large():
    """Automatic policy: a single 2^14 x 8 poly is split, 64 of them are not; both
    give the oracle's words."""
    n, L = 1 << 14, 8
    a = capi.sample_uniform(n, L, 5)
    b = capi.sample_gaussian(n, L, 3.2, 6)
    got, _, _ = _multiply(a, b, n, L, 1, 0)
    np.testing.assert_array_equal(got, orc.polymul(a, b, n, L))


def test_split_policy_rejects():
    lib = capi.load()
    for bad in (1, 6, -2):
        assert lib.ya_split_policy(bad) == 1


@pytest.mark.parametrize("logn", [11, 12, 13, 14])
def test_plan_split_form(logn):
    """A small one-class plan runs in the split form (run and graph replays)
    and matches the oracle; forcing the split form on a larger plan too."""
    rng = np.random.default_rng(logn)
    host = _jobs(rng, 3, [logn], 5)
    say("plan split", logn, [(n, L) for n, L, _, _ in host])
    bt = Batch(host)
    form = ctypes.c_int()
    capi.call("ya_batch_plan_info", bt.plan, None, None, None, ctypes.byref(form))
    assert form.value == 3
    want = [orc.polymul(a, b, n, L) for n, L, a, b in host]
    capi.call("ya_batch_plan_run", bt.plan, None)
    capi.call("ya_device_synchronize")
    for w, got in zip(want, bt.results()):
        np.testing.assert_array_equal(got, w)
    for rep in range(3):
        for _, _, dc in bt.bufs:
            capi.call("ya_memset_zero", dc.ptr, dc.nwords * 8, None)
        capi.call("ya_batch_plan_run_graph", bt.plan, None)
        capi.call("ya_device_synchronize")
        for w, got in zip(want, bt.results()):
            np.testing.assert_array_equal(got, w)
    bt.close()


def test_forced_split_form_on_larger_plan():
    rng = np.random.default_rng(99)
    host = _jobs(rng, 40, [13], 12)
    capi.call("ya_batch_force_form", 3)
    bt = Batch(host)
    form = ctypes.c_int()
    capi.call("ya_batch_plan_info", bt.plan, None, None, None, ctypes.byref(form))
    assert form.value == 3
    capi.call("ya_batch_plan_run", bt.plan, None)
    capi.call("ya_device_synchronize")
    for (n, L, a, b), got in zip(host, bt.results()):
        np.testing.assert_array_equal(got, orc.polymul(a, b, n, L))
    bt.close()
