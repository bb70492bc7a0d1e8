"""The fused steps of the reference example's RLWE flow over many polynomials
(ya_poly_forward_lazy_add_product_many, ya_poly_add_product_backward_many):
per polynomial the oracle's composition of the single steps (example.cpp:41-48,
55-57: operator+=, forward_lazy, add_product_of; add_product_of, backward),
at the fused degrees (2^13, 2^14; an odd last polynomial through the separate
launches) and the others (2^12, 2^16), with and without the addition, pairing
on and off, in place and out of place; overlapping buffers are rejected."""
import numpy as np
import pytest

from _oracle import Oracle, rand_residues
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
O = Oracle(64)

SHAPES = [(4, 3, 3), (4096, 3, 3), (8192, 4, 5), (16384, 2, 2), (65536, 2, 3), (1 << 19, 1, 2)]


def _polys(rng, n, L, P, count, mult=1):
    return np.concatenate([rand_residues(rng, n, L, P, mult) for _ in range(count)])


@pytest.mark.parametrize("with_add", [1, 0])
@pytest.mark.parametrize("n,L,count", SHAPES)
def test_forward_lazy_add_product_many(n, L, count, with_add):
    device()
    rng = np.random.default_rng(n + L + count + with_add)
    P = [O.prime(j) for j in range(L)]
    acc, add, u = (_polys(rng, n, L, P, count) for _ in range(3))
    key = rand_residues(rng, n, L, P)
    acc[:3] = [P[0] - 1, 0, 1]
    add[:3] = [P[0] - 1, P[0] - 1, 0]
    w = n * L
    want = []
    for k in range(count):
        s = slice(k * w, (k + 1) * w)
        x = O.poly_add(acc[s], add[s], n, L) if with_add else acc[s]
        want.append(O.poly_add_product_of(O.poly_forward_lazy(x, n, L), u[s], key, n, L))
    want = np.concatenate(want)
    dadd, du, dkey = (capi.DeviceBuffer.from_host(x) for x in (add, u, key))
    try:
        for pair in (1, 0):
            say("fused forward + add_product_of", n, L, count, "add", with_add, "pair", pair)
            capi.call("ya_pair_forward", pair)
            dacc = capi.DeviceBuffer.from_host(acc)
            capi.call("ya_poly_forward_lazy_add_product_many", dacc.ptr, dadd.ptr if with_add else None, du.ptr,
                      dkey.ptr, n, L, count, None)
            np.testing.assert_array_equal(dacc.download(), want)
            dacc.close()
    finally:
        capi.call("ya_pair_forward", 1)


@pytest.mark.parametrize("in_place", [1, 0])
@pytest.mark.parametrize("n,L,count", SHAPES)
def test_add_product_backward_many(n, L, count, in_place):
    device()
    rng = np.random.default_rng(7 * n + L + count + in_place)
    P = [O.prime(j) for j in range(L)]
    acc, x = (_polys(rng, n, L, P, count) for _ in range(2))
    key = rand_residues(rng, n, L, P)
    acc[:2] = [P[0] - 1, 0]
    x[:2] = [P[0] - 1, P[0] - 1]
    w = n * L
    want = np.concatenate([O.poly_backward(O.poly_add_product_of(acc[k * w:(k + 1) * w], x[k * w:(k + 1) * w],
                                                                 key, n, L), n, L) for k in range(count)])
    say("fused add_product_of + backward", n, L, count, "in place", in_place)
    dacc, dx, dkey = (capi.DeviceBuffer.from_host(v) for v in (acc, x, key))
    dout = dacc if in_place else capi.DeviceBuffer(acc.size)
    capi.call("ya_poly_add_product_backward_many", dout.ptr, dacc.ptr, dx.ptr, dkey.ptr, n, L, count, None)
    np.testing.assert_array_equal(dout.download(), want)
    if not in_place:
        np.testing.assert_array_equal(dacc.download(), acc)


def test_fused_steps_reject_overlaps():
    device()
    n, L, count = 8192, 2, 2
    w = n * L
    big = capi.DeviceBuffer(4 * w * count)
    key = capi.DeviceBuffer(w)
    say("fused steps, overlapping buffers")
    lib = capi.load()
    acc, other = big.ptr, big.ptr + 8 * w * count
    # add / u overlapping acc, key inside acc
    assert lib.ya_poly_forward_lazy_add_product_many(acc, acc + 8 * w, other, key.ptr, n, L, count, None) != 0
    assert lib.ya_poly_forward_lazy_add_product_many(acc, None, acc, key.ptr, n, L, count, None) != 0
    assert lib.ya_poly_forward_lazy_add_product_many(acc, None, other, acc + 8 * w, n, L, count, None) != 0
    # out overlapping x, or acc shifted
    assert lib.ya_poly_add_product_backward_many(acc, other, acc + 8 * w, key.ptr, n, L, count, None) != 0
    assert lib.ya_poly_add_product_backward_many(acc + 8 * w, acc, other, key.ptr, n, L, count, None) != 0
