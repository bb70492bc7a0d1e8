"""The 32-bit moduli build's host side through the C ABI (no device needed):
constants against the generated table, host tables against the oracle."""
import numpy as np
import pytest

from _oracle import Oracle
from yell_amd import capi

O = Oracle(32)


def test_constants():
    lib = capi.load()
    assert lib.ya32_num_moduli() == 291
    for cm in (0, 1, 100, 290):
        assert lib.ya32_prime(cm) == O.prime(cm)
        assert lib.ya32_barrett_const(cm) == O.lib.or_barrett_const(cm)
        assert lib.ya32_root(cm) == O.lib.or_root(cm)
    assert lib.ya32_prime(291) == 0


@pytest.mark.parametrize("logn", [3, 4, 9, 14, 15])
def test_host_tables_match_oracle(logn):
    lib = capi.load()
    n = 1 << logn
    for cm in (0, 290):
        arrs = [np.empty(n, dtype=np.uint32) for _ in range(4)]
        sc = np.empty(4, dtype=np.uint32)
        assert lib.ya32_ntt_table_host(n, cm, *[capi.ptr(a) for a in arrs], capi.ptr(sc)) == 0
        for got, want in zip(arrs + [sc], O.tables(n, cm)):
            np.testing.assert_array_equal(got, want)


def test_rejects():
    lib = capi.load()
    buf = np.zeros(64, dtype=np.uint32)
    p = capi.ptr(buf)
    assert lib.ya32_ntt_table_host(1 << 16, 0, p, p, p, p, p) == 1    # beyond 2^15
    assert lib.ya32_ntt_table_host(4, 0, p, p, p, p, p) == 1          # below 2^3
    assert lib.ya32_ntt_table_host(8, 291, p, p, p, p, p) == 1        # prime index
    assert lib.ya32_poly_multiply(None, None, None, 8, 1, None) == 1
