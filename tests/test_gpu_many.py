"""The pointwise operations over many polynomials (ya_poly_*_many): each
polynomial against the oracle's single form (details/poly.hpp:119-158,
ops.hpp:7-41), with b per polynomial and b shared, in place, at the lazy
ranges' edges; a shared b that overlaps out is rejected."""
import numpy as np
import pytest

from _oracle import Oracle, rand_residues
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
O = Oracle(64)

OPS = {"mul": "poly_mul", "add": "poly_add", "sub": "poly_sub", "add_product_of": "poly_add_product_of"}


@pytest.mark.parametrize("op", sorted(OPS))
@pytest.mark.parametrize("shared", [0, 1])
@pytest.mark.parametrize("n,L,count", [(8, 1, 3), (4096, 4, 7), (16384, 8, 20)])
def test_pointwise_many(op, shared, n, L, count):
    say("pointwise many", op, shared, n, L, count)
    device()
    rng = np.random.default_rng(n + L + count)
    P = [O.prime(j) for j in range(L)]
    mult = 4 if op == "add_product_of" else 1   # the accumulator may hold forward_lazy words
    a = np.concatenate([rand_residues(rng, n, L, P) for _ in range(count)])
    b = rand_residues(rng, n, L, P) if shared else np.concatenate([rand_residues(rng, n, L, P) for _ in range(count)])
    acc = np.concatenate([rand_residues(rng, n, L, P, mult) for _ in range(count)])
    a[:2] = [P[0] - 1, 0]
    da, db, dc = (capi.DeviceBuffer.from_host(x) for x in (a, b, acc))
    if op == "add_product_of":
        capi.call("ya_poly_add_product_of_many", dc.ptr, da.ptr, db.ptr, n, L, count, shared, None)
    else:
        capi.call("ya_poly_%s_many" % op, dc.ptr, da.ptr, db.ptr, n, L, count, shared, None)
    got = dc.download().reshape(count, n * L)
    w = n * L
    for k in range(count):
        bk = b if shared else b[k * w:(k + 1) * w]
        ak = a[k * w:(k + 1) * w]
        if op == "add_product_of":
            want = O.poly_add_product_of(acc[k * w:(k + 1) * w], ak, bk, n, L)
        else:
            want = getattr(O, OPS[op])(ak, bk, n, L)
        np.testing.assert_array_equal(got[k], want, err_msg="polynomial %d" % k)
    for d in (da, db, dc):
        d.close()


def test_in_place_and_negate_many():
    device()
    n, L, count = 1024, 3, 4
    rng = np.random.default_rng(3)
    P = [O.prime(j) for j in range(L)]
    a = np.concatenate([rand_residues(rng, n, L, P) for _ in range(count)])
    a[0] = 0
    b = rand_residues(rng, n, L, P)
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    capi.call("ya_poly_add_many", da.ptr, da.ptr, db.ptr, n, L, count, 1, None)   # a_k += b
    capi.call("ya_poly_negate_many", da.ptr, n, L, count, None)
    got = da.download().reshape(count, n * L)
    w = n * L
    for k in range(count):
        want = O.poly_negate(O.poly_add(a[k * w:(k + 1) * w], b, n, L), n, L)
        np.testing.assert_array_equal(got[k], want)
    lib = capi.load()
    assert lib.ya_poly_add_many(da.ptr, da.ptr, da.ptr + 8 * w, n, L, count, 1, None) == 1   # shared b inside out
    assert lib.ya_poly_mul_many(da.ptr, da.ptr, db.ptr, n, L, 0, 0, None) == 1              # count 0
    da.close()
    db.close()


def test_pointwise_many_chunks():
    """count * nmoduli beyond one launch's 65,535 grid rows: 70 polynomials of
    1000 residues (the launches are cut at 65 polynomials)."""
    device()
    n, L, count = 8, 1000, 70
    rng = np.random.default_rng(70)
    P = [O.prime(j) for j in range(L)]
    a = np.concatenate([rand_residues(rng, n, L, P) for _ in range(count)])
    b = rand_residues(rng, n, L, P)
    acc = np.concatenate([rand_residues(rng, n, L, P, 4) for _ in range(count)])
    da, db, dc = (capi.DeviceBuffer.from_host(x) for x in (a, b, acc))
    say("pointwise many chunks")
    capi.call("ya_poly_add_product_of_many", dc.ptr, da.ptr, db.ptr, n, L, count, 1, None)
    got = dc.download().reshape(count, n * L)
    w = n * L
    for k in (0, 64, 65, count - 1):
        np.testing.assert_array_equal(got[k], O.poly_add_product_of(acc[k * w:(k + 1) * w], a[k * w:(k + 1) * w], b, n, L))
    for d in (da, db, dc):
        d.close()


def test_pointwise_range_edges_everywhere():
    """Every word at the range edges, so that every lane takes the largest
    128-bit products (the middle column's carry out of bit 63) and Barrett
    quotients: a = b = p - 1, the accumulator at 4p - 1 (a forward_lazy
    word), and the mixed edges 0, 1, p - 1 against each other."""
    device()
    n, L, count = 4096, 24, 2
    P = [O.prime(j) for j in range(L)]
    top = np.concatenate([np.full(n, p - 1, dtype=np.uint64) for p in P])
    acc4 = np.concatenate([np.full(n, 4 * p - 1, dtype=np.uint64) for p in P])
    mixed = np.tile(np.array([0, 1, 2, 3], dtype=np.uint64), n * L // 4)
    mixed = np.concatenate([np.where(mixed[r * n:(r + 1) * n] == 3, P[r] - 1, mixed[r * n:(r + 1) * n]) for r in range(L)])
    cases = [("mul", top, top, None), ("add_product_of", top, top, acc4), ("mul", mixed, top, None),
             ("add_product_of", top, mixed, acc4)]
    for op, a1, b1, acc1 in cases:
        say("range edges", op)
        a = np.concatenate([a1] * count)
        da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b1)
        dc = capi.DeviceBuffer.from_host(np.concatenate([acc1] * count) if acc1 is not None else np.zeros_like(a))
        capi.call("ya_poly_%s_many" % op, dc.ptr, da.ptr, db.ptr, n, L, count, 1, None)
        got = dc.download().reshape(count, n * L)
        want = O.poly_add_product_of(acc1, a1, b1, n, L) if acc1 is not None else O.poly_mul(a1, b1, n, L)
        for k in range(count):
            np.testing.assert_array_equal(got[k], want, err_msg="%s polynomial %d" % (op, k))
        for d in (da, db, dc):
            d.close()
