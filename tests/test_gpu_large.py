"""Large-ring multiply-accumulate (N = 2^16, L = 24, 8 pairs) and the
many-polynomial transforms, bit-exact against the oracle."""
import ctypes

import numpy as np
import pytest

from _oracle import Oracle
from gpu_util import device, rand_bound, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
orc = Oracle()
P = [orc.prime(i) for i in range(32)]


@pytest.mark.parametrize("logn", [2, 10, 14, 15, 16, 18, 19])
def test_many_transforms(logn):
    device()
    n, L, count = 1 << logn, (3 if logn <= 16 else 1), 3
    rng = np.random.default_rng(logn)
    x = np.concatenate([rand_bound(rng, n, L, 4, P) for _ in range(count)])
    d = capi.DeviceBuffer.from_host(x)
    say("forward_lazy_many", logn)
    capi.call("ya_poly_forward_lazy_many", d.ptr, n, L, count, None)
    got = d.download()
    for k in range(count):
        s = slice(k * n * L, (k + 1) * n * L)
        np.testing.assert_array_equal(got[s], orc.poly_forward_lazy(x[s], n, L))
    y = np.concatenate([rand_bound(rng, n, L, 2, P) for _ in range(count)])
    d.upload(y)
    capi.call("ya_poly_backward_many", d.ptr, n, L, count, None)
    got = d.download()
    for k in range(count):
        s = slice(k * n * L, (k + 1) * n * L)
        np.testing.assert_array_equal(got[s], orc.poly_backward(y[s], n, L))


@pytest.mark.parametrize("logn", [12, 13, 14])
@pytest.mark.parametrize("le", [3, 4])
def test_paired_forward_many(logn, le):
    """The paired forward (two same-prime residues per workgroup) against the
    oracle per polynomial, for 2, 3 and 5 polynomials (the odd last one runs
    alone), lazy and corrected, inputs up to 4p - 1; and the same words with
    pairing off."""
    if le == 3 and logn == 14:
        pytest.skip("radix-8 passes stop at 2^13")
    lib = device()
    n, L = 1 << logn, 3
    old = lib.ya_tune_get(0, logn)
    assert lib.ya_tune_set(0, logn, le) == 0
    try:
        rng = np.random.default_rng(logn * 10 + le)
        for count in (2, 3, 5):
            x = np.concatenate([rand_bound(rng, n, L, 4, P) for _ in range(count)])
            x[:3] = [4 * P[0] - 1, 0, 2 * P[0]]
            for fn, ref in (("ya_poly_forward_lazy_many", orc.poly_forward_lazy),
                            ("ya_poly_forward_many", orc.poly_forward)):
                say("paired forward", logn, le, count, fn)
                outs = []
                for pair in (1, 0):
                    capi.call("ya_pair_forward", pair)
                    d = capi.DeviceBuffer.from_host(x)
                    capi.call(fn, d.ptr, n, L, count, None)
                    outs.append(d.download())
                    d.close()
                capi.call("ya_pair_forward", 1)
                np.testing.assert_array_equal(outs[0], outs[1])
                for k in range(count):
                    s = slice(k * n * L, (k + 1) * n * L)
                    np.testing.assert_array_equal(outs[0][s], ref(x[s], n, L), err_msg="poly %d" % k)
    finally:
        lib.ya_tune_set(0, logn, old)
        capi.call("ya_pair_forward", 1)


@pytest.mark.parametrize("logn", [15, 16, 17, 18, 19])
def test_paired_forward_large_rings(logn):
    """Above 2^14 the paired forward runs the top layers per residue, then
    block b of residue r of polynomials 2q and 2q+1 in one workgroup: the
    oracle's words per polynomial for 2 and 3 polynomials (the odd last one
    alone), lazy and corrected, inputs up to 4p - 1, and the same words with
    pairing off."""
    device()
    n, L = 1 << logn, (2 if logn <= 17 else 1)
    rng = np.random.default_rng(logn)
    try:
        for count in (2, 3):
            x = np.concatenate([rand_bound(rng, n, L, 4, P) for _ in range(count)])
            x[:3] = [4 * P[0] - 1, 0, 2 * P[0]]
            for fn, ref in (("ya_poly_forward_lazy_many", orc.poly_forward_lazy),
                            ("ya_poly_forward_many", orc.poly_forward)):
                say("paired forward, large ring", logn, count, fn)
                outs = []
                for pair in (1, 0):
                    capi.call("ya_pair_forward", pair)
                    d = capi.DeviceBuffer.from_host(x)
                    capi.call(fn, d.ptr, n, L, count, None)
                    outs.append(d.download())
                    d.close()
                capi.call("ya_pair_forward", 1)
                np.testing.assert_array_equal(outs[0], outs[1])
                for k in range(count):
                    s = slice(k * n * L, (k + 1) * n * L)
                    np.testing.assert_array_equal(outs[0][s], ref(x[s], n, L), err_msg="poly %d" % k)
    finally:
        capi.call("ya_pair_forward", 1)


@pytest.mark.parametrize("npairs", [1, 3, 17])
def test_add_products(npairs):
    device()
    n, L = 1 << 11, 4
    rng = np.random.default_rng(npairs)
    acc = rand_bound(rng, n, L, 1, P)
    a = [rand_bound(rng, n, L, 1, P) for _ in range(npairs)]
    b = [rand_bound(rng, n, L, 1, P) for _ in range(npairs)]
    da = [capi.DeviceBuffer.from_host(v) for v in a]
    db = [capi.DeviceBuffer.from_host(v) for v in b]
    dacc = capi.DeviceBuffer.from_host(acc)
    pa = (ctypes.c_void_p * npairs)(*[x.ptr for x in da])
    pb = (ctypes.c_void_p * npairs)(*[x.ptr for x in db])
    say("add_products", npairs)
    capi.call("ya_poly_add_products", dacc.ptr, pa, pb, npairs, n, L, None)
    want = acc
    for j in range(npairs):
        want = orc.poly_add_product_of(want, a[j], b[j], n, L)
    np.testing.assert_array_equal(dacc.download(), want)


def test_c4_large_ring_mac():
    """sum_{j<8} a_j * b_j at N = 2^16, L = 24: 16 forward NTTs, the MAC, one
    inverse NTT (the reference sequence: forward, add_product_of x 8, backward)."""
    device()
    n, L, pairs = 1 << 16, 24, 8
    ops = [capi.sample_uniform(n, L, 4000 + k) for k in range(2 * pairs)]
    dall = capi.DeviceBuffer.from_host(np.concatenate(ops))
    w = n * L
    dacc = capi.DeviceBuffer(w)
    capi.call("ya_memset_zero", dacc.ptr, w * 8, None)
    say("2^16 x 24: forward x16")
    capi.call("ya_poly_forward_many", dall.ptr, n, L, 2 * pairs, None)
    pa = (ctypes.c_void_p * pairs)(*[dall.ptr + 8 * w * (2 * j) for j in range(pairs)])
    pb = (ctypes.c_void_p * pairs)(*[dall.ptr + 8 * w * (2 * j + 1) for j in range(pairs)])
    capi.call("ya_poly_add_products", dacc.ptr, pa, pb, pairs, n, L, None)
    capi.call("ya_poly_backward", dacc.ptr, n, L, None)
    got = dacc.download()
    say("2^16 x 24: oracle")
    acc = np.zeros(w, dtype=np.uint64)
    for j in range(pairs):
        fa = orc.poly_forward(ops[2 * j], n, L)
        fb = orc.poly_forward(ops[2 * j + 1], n, L)
        acc = orc.poly_add_product_of(acc, fa, fb, n, L)
    np.testing.assert_array_equal(got, orc.poly_backward(acc, n, L))


def _mac_want(acc, a, b, n, L):
    want = acc
    for x, y in zip(a, b):
        want = orc.poly_add_product_of(want, orc.poly_forward(x, n, L), orc.poly_forward(y, n, L), n, L)
    return orc.poly_backward(want, n, L)


@pytest.mark.parametrize("logn,L,npairs,lb", [(12, 2, 3, 14), (13, 3, 3, 14), (14, 2, 17, 14), (16, 5, 4, 14),
                                              (16, 5, 4, 13), (16, 5, 4, 12), (15, 3, 17, 12), (17, 2, 2, 12),
                                              (18, 2, 2, 13), (19, 1, 2, 14)])
def test_add_products_backward(logn, L, npairs, lb):
    """ya_poly_add_products_backward = the reference's sequence a_j.forward();
    b_j.forward(); acc.add_product_of(a_j, b_j) for every pair, then
    acc.backward(), from a non-zero NTT-domain acc: below 2^13 as separate
    launches, from 2^13 the pair-product pass (block 2^lb) and the summing
    inverse, 17 pairs in two chunks."""
    lib = device()
    n = 1 << logn
    rng = np.random.default_rng(logn * 100 + npairs)
    acc = rand_bound(rng, n, L, 1, P)
    a = [rand_bound(rng, n, L, 1, P) for _ in range(npairs)]
    b = [rand_bound(rng, n, L, 1, P) for _ in range(npairs)]
    a[0][:3] = [P[0] - 1, 0, P[0] - 1]
    b[0][:3] = [P[0] - 1, P[0] - 1, 0]
    old = lib.ya_tune_get(4, logn)
    capi.call("ya_tune_set", 4, logn, lb)
    try:
        da = [capi.DeviceBuffer.from_host(v) for v in a]
        db = [capi.DeviceBuffer.from_host(v) for v in b]
        dacc = capi.DeviceBuffer.from_host(acc)
        pa = (ctypes.c_void_p * npairs)(*[x.ptr for x in da])
        pb = (ctypes.c_void_p * npairs)(*[x.ptr for x in db])
        say("add_products_backward", logn, L, npairs, lb)
        capi.call("ya_poly_add_products_backward", dacc.ptr, pa, pb, npairs, n, L, None)
        np.testing.assert_array_equal(dacc.download(), _mac_want(acc, a, b, n, L))
    finally:
        capi.call("ya_tune_set", 4, logn, old if old in (12, 13, 14) else 14)


def test_add_products_backward_c4_and_rejects_overlap():
    """The large-ring job (N = 2^16, L = 24, 8 pairs, acc = 0) in one call
    against the oracle's sequence; overlapping operands are refused."""
    lib = device()
    n, L, pairs = 1 << 16, 24, 8
    w = n * L
    ops = [capi.sample_uniform(n, L, 6000 + k) for k in range(2 * pairs)]
    dall = capi.DeviceBuffer.from_host(np.concatenate(ops))
    dacc = capi.DeviceBuffer(w)
    capi.call("ya_memset_zero", dacc.ptr, w * 8, None)
    pa = (ctypes.c_void_p * pairs)(*[dall.ptr + 8 * w * (2 * j) for j in range(pairs)])
    pb = (ctypes.c_void_p * pairs)(*[dall.ptr + 8 * w * (2 * j + 1) for j in range(pairs)])
    say("add_products_backward 2^16 x 24, 8 pairs")
    capi.call("ya_poly_add_products_backward", dacc.ptr, pa, pb, pairs, n, L, None)
    got = dacc.download()
    np.testing.assert_array_equal(got, _mac_want(np.zeros(w, dtype=np.uint64), ops[0::2], ops[1::2], n, L))
    pb[3] = pa[2]
    assert lib.ya_poly_add_products_backward(dacc.ptr, pa, pb, pairs, n, L, None) == 1
    pb[3] = dacc.ptr
    assert lib.ya_poly_add_products_backward(dacc.ptr, pa, pb, pairs, n, L, None) == 1


@pytest.mark.parametrize("logn,L,npairs,lb", [(12, 2, 3, 14), (13, 3, 1, 14), (14, 2, 17, 14), (16, 5, 4, 13),
                                              (16, 24, 8, 13), (16, 24, 8, 12), (16, 3, 33, 12), (19, 1, 2, 14)])
def test_sum_products_backward(logn, L, npairs, lb):
    """ya_poly_sum_products_backward = the reference's sequence on a fresh
    (zero-filled) accumulator: below 2^13 the first pair's product written by
    the pointwise multiply, then the MAC; from 2^13 the summing inverse that
    never reads the output (which holds garbage words beforehand), 17 and 33
    pairs in two and three chunks (the first chunk's first product copied)."""
    lib = device()
    n = 1 << logn
    rng = np.random.default_rng(logn * 1000 + npairs * 10 + L)
    a = [rand_bound(rng, n, L, 1, P) for _ in range(npairs)]
    b = [rand_bound(rng, n, L, 1, P) for _ in range(npairs)]
    a[0][:3] = [P[0] - 1, 0, P[0] - 1]
    b[0][:3] = [P[0] - 1, P[0] - 1, 0]
    old = lib.ya_tune_get(4, logn)
    capi.call("ya_tune_set", 4, logn, lb)
    try:
        da = [capi.DeviceBuffer.from_host(v) for v in a]
        db = [capi.DeviceBuffer.from_host(v) for v in b]
        dout = capi.DeviceBuffer.from_host(rand_bound(rng, n, L, 1, P) | np.uint64(1 << 63))
        pa = (ctypes.c_void_p * npairs)(*[x.ptr for x in da])
        pb = (ctypes.c_void_p * npairs)(*[x.ptr for x in db])
        say("sum_products_backward", logn, L, npairs, lb)
        capi.call("ya_poly_sum_products_backward", dout.ptr, pa, pb, npairs, n, L, None)
        np.testing.assert_array_equal(dout.download(), _mac_want(np.zeros(n * L, dtype=np.uint64), a, b, n, L))
        assert lib.ya_poly_sum_products_backward(dout.ptr, pa, pb, 0, n, L, None) == 1
        pb[0] = dout.ptr
        assert lib.ya_poly_sum_products_backward(dout.ptr, pa, pb, npairs, n, L, None) == 1
    finally:
        capi.call("ya_tune_set", 4, logn, old if old in (12, 13, 14) else 14)


def test_autotune_mac_times_every_accepted_block_size():
    """ya_autotune_mac times the job with each accepted pair-pass block size
    (2^14, 2^13, 2^12 at 2^16; 2^12 refused at 2^18) and keeps one of them;
    the job then still gives the oracle's words."""
    lib = device()
    saved = {lg: lib.ya_tune_get(4, lg) for lg in (16, 18)}
    try:
        rep = (ctypes.c_double * 3)()
        say("autotune_mac 2^16 x 3, 2 pairs")
        capi.call("ya_autotune_mac", 1 << 16, 3, 2, rep)
        assert all(rep[k] > 0 for k in range(3))
        assert lib.ya_tune_get(4, 16) in (12, 13, 14)
        say("autotune_mac 2^18 x 1, 1 pair")
        capi.call("ya_autotune_mac", 1 << 18, 1, 1, rep)
        assert rep[0] > 0 and rep[1] > 0 and rep[2] == -1.0
        assert lib.ya_tune_get(4, 18) in (13, 14)
        n, L, npairs = 1 << 16, 3, 2
        rng = np.random.default_rng(16)
        a = [rand_bound(rng, n, L, 1, P) for _ in range(npairs)]
        b = [rand_bound(rng, n, L, 1, P) for _ in range(npairs)]
        da = [capi.DeviceBuffer.from_host(v) for v in a]
        db = [capi.DeviceBuffer.from_host(v) for v in b]
        dout = capi.DeviceBuffer(n * L)
        pa = (ctypes.c_void_p * npairs)(*[x.ptr for x in da])
        pb = (ctypes.c_void_p * npairs)(*[x.ptr for x in db])
        capi.call("ya_poly_sum_products_backward", dout.ptr, pa, pb, npairs, n, L, None)
        np.testing.assert_array_equal(dout.download(), _mac_want(np.zeros(n * L, dtype=np.uint64), a, b, n, L))
    finally:
        for lg, v in saved.items():
            capi.call("ya_tune_set", 4, lg, v)
