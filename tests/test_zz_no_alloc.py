"""ya_poly_multiply_noalloc (the entry point that allocates nothing) against
ya_poly_multiply, bit for bit, through the C ABI: two calls in a row on each
path that uses scratch (the split launch of few residues, and a ring above
2^14), the second with a smaller input, plus the oracle on the small shapes
and the refusals (scratch too small, tables not resident).
"""
import numpy as np
import pytest

from _oracle import Oracle
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
orc = Oracle()


def _operands(n, L, seed):
    return capi.sample_uniform(n, L, seed), capi.sample_gaussian(n, L, 3.2, seed + 1)


def _both(n, L, seed):
    """(noalloc result, ya_poly_multiply result, a, b) for one shape."""
    a, b = _operands(n, L, seed)
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    dc, dr = capi.DeviceBuffer(n * L), capi.DeviceBuffer(n * L)
    capi.call("ya_poly_multiply_noalloc", dc.ptr, da.ptr, db.ptr, n, L, None)
    capi.call("ya_stream_synchronize", None)
    got = dc.download()
    capi.call("ya_poly_multiply", dr.ptr, da.ptr, db.ptr, n, L, None)
    want = dr.download()
    np.testing.assert_array_equal(a, da.download())   # operands untouched
    np.testing.assert_array_equal(b, db.download())
    for buf in (da, db, dc, dr):
        buf.close()
    return got, want, a, b


@pytest.mark.parametrize("shapes", [
    # the split launch (few residues of 2^11..2^14): 2^14 x 8, then 2^12 x 3
    [(1 << 14, 8), (1 << 12, 3)],
    # the large-ring path: 2^16 x 4, then 2^15 x 2
    [(1 << 16, 4), (1 << 15, 2)],
    # the fused kernel (no scratch), then the split launch
    [(1 << 12, 160), (1 << 11, 1)],
])
def test_noalloc_matches_poly_multiply(shapes):
    device()
    capi.call("ya_split_policy", 0)
    for n, L in shapes:
        capi.call("ya_ntt_init_tables", n, L)
    for k, (n, L) in enumerate(shapes):
        say("noalloc poly-mult N=%d L=%d call %d" % (n, L, k))
        got, want, a, b = _both(n, L, 100 + k)
        np.testing.assert_array_equal(got, want)
        if n * L <= (1 << 14):
            np.testing.assert_array_equal(got, orc.polymul(a, b, n, L))


def test_noalloc_refusals():
    lib = device()
    n = 1 << 17
    capi.call("ya_ntt_init_tables", n, 17)
    big = capi.DeviceBuffer(3 * n * 17)
    a, b, c = big.ptr, big.ptr + n * 17 * 8, big.ptr + 2 * n * 17 * 8
    say("noalloc refusal: N=2^17 L=17 needs more scratch than the static array")
    assert lib.ya_poly_multiply_noalloc(c, a, b, n, 17, None) == 1
    assert b"scratch" in lib.ya_last_error()
    # tables never built for this shape: refused instead of built
    say("noalloc refusal: tables not resident")
    assert lib.ya_poly_multiply_noalloc(c, a, b, 1 << 11, 777, None) == 1
    assert b"ya_ntt_init_tables" in lib.ya_last_error()
    big.close()
