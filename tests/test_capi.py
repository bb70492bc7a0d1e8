"""CPU-only checks of the C ABI: the library loads and exports every declared
symbol, the host table builder and samplers behave, bad arguments are
rejected with status codes.  No call here launches GPU work."""
import ctypes
import glob
import os

import numpy as np
import pytest

from _oracle import Oracle
from yell_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
orc = Oracle()


def test_exports_every_declared_symbol():
    lib = capi.load()
    names = capi.declared_symbols(sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_constants():
    lib = capi.load()
    assert lib.ya_num_moduli() == 1000
    for cm in (0, 1, 999):
        assert lib.ya_prime(cm) == orc.prime(cm)
        assert lib.ya_barrett_const(cm) == orc.lib.or_barrett_const(cm)
        assert lib.ya_root(cm) == orc.lib.or_root(cm)
    assert lib.ya_prime(1000) == 0


@pytest.mark.parametrize("logn", [3, 10, 14, 17])
def test_host_tables_match_oracle(logn):
    n = 1 << logn
    for cm in (0, 13):
        for a, b in zip(capi.table_host(n, cm), orc.tables(n, cm)):
            np.testing.assert_array_equal(a, b)


def test_uniform_sampler():
    n, L = 4096, 3
    a = capi.sample_uniform(n, L, 1)
    np.testing.assert_array_equal(a, capi.sample_uniform(n, L, 1))
    assert not np.array_equal(a, capi.sample_uniform(n, L, 2))
    for r in range(L):
        res = a[r * n:(r + 1) * n]
        assert res.max() < capi.primes(L)[r]
        # roughly uniform over [0, p): mean within 5 sigma of p/2
        m = res.astype(np.float64).mean() / float(capi.primes(L)[r])
        assert abs(m - 0.5) < 5 * (1 / np.sqrt(12 * n))


def test_gaussian_sampler():
    n, L, sigma = 1 << 16, 2, 3.2
    g = capi.sample_gaussian(n, L, sigma, 9)
    P = capi.primes(L)
    signed = [np.where(g[r * n:(r + 1) * n] > P[r] // 2,
                       g[r * n:(r + 1) * n].astype(np.int64) - np.int64(P[r]),
                       g[r * n:(r + 1) * n].astype(np.int64)) for r in range(L)]
    np.testing.assert_array_equal(signed[0], signed[1])   # one sample shared by residues
    s = signed[0].astype(np.float64)
    assert abs(s.mean()) < 5 * sigma / np.sqrt(n)
    assert abs(s.std() - sigma) < 0.05 * sigma
    assert np.abs(s).max() <= np.ceil(13 * sigma)


def test_invalid_arguments_rejected():
    lib = capi.load()
    buf = np.zeros(16, dtype=np.uint64)
    assert lib.ya_sample_uniform(capi.ptr(buf), 12, 1, 0) == 1        # not a power of two
    assert lib.ya_sample_uniform(capi.ptr(buf), 2, 1, 0) == 1         # below 2^2
    assert lib.ya_sample_uniform(capi.ptr(buf), 8, 1001, 0) == 1      # too many moduli
    assert lib.ya_sample_uniform(None, 8, 1, 0) == 1
    assert lib.ya_sample_gaussian(capi.ptr(buf), 8, 1, -1.0, 0) == 1
    assert b"degree" in lib.ya_last_error() or b"sigma" in lib.ya_last_error()
    assert lib.ya_poly_forward(None, 1 << 20, 1, None) == 1           # 2^20 unsupported (N < 2^20)
    assert b"degree" in lib.ya_last_error()
    assert lib.ya_poly_forward(None, 2, 1, None) == 1                 # 2 unsupported
    assert lib.ya_ntt_table_host(8, 1000, capi.ptr(buf), capi.ptr(buf), capi.ptr(buf),
                                 capi.ptr(buf), capi.ptr(buf)) == 1
    assert lib.ya_status_string(8) == b"tables do not match"


def test_tune_table_defaults_and_bounds():
    """The launch-shape table is host state (no GPU call): its defaults and the
    ranges ya_tune_set accepts for the pair-pass block size (kind 4) and the
    staggered start (kind 5); a refused entry leaves the table unchanged."""
    lib = capi.load()
    assert lib.ya_tune_get(4, 16) == 12 and lib.ya_tune_get(4, 18) == 13 and lib.ya_tune_get(4, 19) == 14
    assert lib.ya_tune_get(5, 14) == 2 and lib.ya_tune_get(5, 15) == 6 and lib.ya_tune_get(5, 13) == 0
    assert lib.ya_tune_set(4, 18, 12) != 0 and lib.ya_tune_set(4, 19, 13) != 0 and lib.ya_tune_set(4, 16, 11) != 0
    assert lib.ya_tune_get(4, 18) == 13
    assert lib.ya_tune_set(5, 14, 17) != 0 and lib.ya_tune_set(5, 14, -1) != 0
    assert lib.ya_tune_get(5, 14) == 2
    assert lib.ya_tune_set(5, 12, 16) == 0 and lib.ya_tune_get(5, 12) == 16
    assert lib.ya_tune_set(5, 12, 0) == 0 and lib.ya_tune_get(5, 12) == 0
    assert lib.ya_tune_get(6, 14) == -1


def test_autotune_mac_rejects_shapes_before_touching_the_device():
    """ya_autotune_mac refuses a ring of one block, no pairs, more pairs than a
    pass takes, and more moduli than the pair pass takes, before any GPU call."""
    lib = capi.load()
    assert lib.ya_autotune_mac(1 << 14, 3, 2, None) != 0
    assert lib.ya_autotune_mac(1 << 16, 3, 0, None) != 0
    assert lib.ya_autotune_mac(1 << 16, 3, 17, None) != 0
    assert lib.ya_autotune_mac(1 << 16, 65536, 2, None) != 0
    assert lib.ya_autotune_mac(3 << 14, 3, 2, None) != 0
