"""The 32-bit moduli build on the GPU (ya32_*): every entry point bit-exact
against the 32-bit oracle, the reference's 32-bit golden vectors, and at the
lazy ranges' maximum inputs."""
import os

import numpy as np
import pytest

from _oracle import Oracle
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
O = Oracle(32)
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ntt32_ref.npz"))


def rand_res(rng, n, L, mult=1):
    return np.concatenate([rng.integers(0, mult * O.prime(j), size=n, dtype=np.uint32) for j in range(L)])


def run1(fn, x, *args):
    device()
    d = capi.DeviceBuffer.from_host(np.ascontiguousarray(x, dtype=np.uint32))
    capi.call(fn, d.ptr, *args, None)
    out = d.download()
    d.close()
    return out


@pytest.mark.parametrize("logn", list(range(3, 16)))
def test_single_residue_transforms(logn):
    n = 1 << logn
    say("ntt32", logn)
    rng = np.random.default_rng(logn)
    for cm in (0, 290):
        p = O.prime(cm)
        x4 = rng.integers(0, 4 * p, size=n, dtype=np.uint32)
        x4[:3] = [4 * p - 1, 0, 2 * p]
        x2 = rng.integers(0, 2 * p, size=n, dtype=np.uint32)
        x2[:2] = [2 * p - 1, 0]
        np.testing.assert_array_equal(run1("ya32_ntt_forward_lazy", x4, n, cm), O.forward_lazy(x4, n, cm))
        np.testing.assert_array_equal(run1("ya32_ntt_forward", x4, n, cm), O.forward(x4, n, cm))
        np.testing.assert_array_equal(run1("ya32_ntt_backward", x2, n, cm), O.backward(x2, n, cm))


@pytest.mark.parametrize("key", sorted({k[len("fwd_in_"):] for k in GOLD.files if k.startswith("fwd_in_")}))
def test_transforms_golden(key):
    n, cm = (int(x[1:] if x[0] == "n" else x[2:]) for x in key.split("_"))
    np.testing.assert_array_equal(run1("ya32_ntt_forward_lazy", GOLD["fwd_in_" + key], n, cm),
                                  GOLD["fwd_lazy_out_" + key])
    np.testing.assert_array_equal(run1("ya32_ntt_backward", GOLD["bwd_in_" + key], n, cm), GOLD["bwd_out_" + key])


@pytest.mark.parametrize("n,L", [(1 << 14, 16), (1 << 15, 3), (1 << 10, 291), (8, 5)])
def test_poly_transforms_and_ops(n, L):
    say("poly32", n, L)
    rng = np.random.default_rng(n + L)
    x = rand_res(rng, n, L, 4)
    np.testing.assert_array_equal(run1("ya32_poly_forward_lazy", x, n, L), O.poly_forward_lazy(x, n, L))
    np.testing.assert_array_equal(run1("ya32_poly_forward", x, n, L), O.poly_forward(x, n, L))
    y = rand_res(rng, n, L, 2)
    np.testing.assert_array_equal(run1("ya32_poly_backward", y, n, L), O.poly_backward(y, n, L))
    a, b, acc = rand_res(rng, n, L), rand_res(rng, n, L), rand_res(rng, n, L)
    a[:4] = 0
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    for fn, want in (("ya32_poly_mul", O.poly_mul(a, b, n, L)), ("ya32_poly_add", O.poly_add(a, b, n, L)),
                     ("ya32_poly_sub", O.poly_sub(a, b, n, L))):
        dc = capi.DeviceBuffer(n * L, np.uint32)
        capi.call(fn, dc.ptr, da.ptr, db.ptr, n, L, None)
        np.testing.assert_array_equal(dc.download(), want)
        dc.close()
    np.testing.assert_array_equal(run1("ya32_poly_negate", a, n, L), O.poly_negate(a, n, L))
    dacc = capi.DeviceBuffer.from_host(acc)
    capi.call("ya32_poly_add_product_of", dacc.ptr, da.ptr, db.ptr, n, L, None)
    np.testing.assert_array_equal(dacc.download(), O.poly_add_product_of(acc, a, b, n, L))
    for d in (da, db, dacc):
        d.close()


@pytest.mark.parametrize("n,L,count", [(1 << 14, 16, 3), (1 << 15, 4, 2), (1 << 12, 291, 1), (8, 1, 4),
                                       (1 << 13, 7, 5)])
def test_polymul_many(n, L, count):
    say("polymul32", n, L, count)
    rng = np.random.default_rng(7 * n + L)
    a = np.concatenate([rand_res(rng, n, L) for _ in range(count)])
    b = np.concatenate([rand_res(rng, n, L) for _ in range(count)])
    b[:n] = np.concatenate([np.full(n, O.prime(0) - 1, dtype=np.uint32)])
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    dc = capi.DeviceBuffer(a.size, np.uint32)
    capi.call("ya32_poly_multiply_many", dc.ptr, da.ptr, db.ptr, n, L, count, None)
    got = dc.download()
    w = n * L
    for k in range(count):
        np.testing.assert_array_equal(got[k * w:(k + 1) * w], O.polymul(a[k * w:(k + 1) * w], b[k * w:(k + 1) * w], n, L))
    np.testing.assert_array_equal(da.download(), a)
    for d in (da, db, dc):
        d.close()


@pytest.mark.parametrize("key", sorted({k[5:] for k in GOLD.files if k.startswith("pm_c_")}))
def test_polymul_golden(key):
    n, L = (int(x[1:]) for x in key.split("_"))
    device()
    da = capi.DeviceBuffer.from_host(GOLD["pm_a_" + key])
    db = capi.DeviceBuffer.from_host(GOLD["pm_b_" + key])
    dc = capi.DeviceBuffer(n * L, np.uint32)
    capi.call("ya32_poly_multiply", dc.ptr, da.ptr, db.ptr, n, L, None)
    np.testing.assert_array_equal(dc.download(), GOLD["pm_c_" + key])
    for d in (da, db, dc):
        d.close()


def test_schoolbook_c1_like():
    """N = 4096, one prime: the poly-mult is the negacyclic convolution."""
    n = 4096
    rng = np.random.default_rng(41)
    a, b = rand_res(rng, n, 1), rand_res(rng, n, 1)
    device()
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    dc = capi.DeviceBuffer(n, np.uint32)
    capi.call("ya32_poly_multiply", dc.ptr, da.ptr, db.ptr, n, 1, None)
    np.testing.assert_array_equal(dc.download(), O.schoolbook(a, b, n, 0))
    for d in (da, db, dc):
        d.close()
