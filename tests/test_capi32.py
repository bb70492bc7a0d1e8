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


def test_samplers_32():
    """Uniform words below each 30-bit prime; Gaussian / ZO / HWT shared by all
    residues with -1 stored as p - 1 (the 64-bit samplers' rules)."""
    lib = capi.load()
    n, L = 4096, 3
    ps = [O.prime(j) for j in range(L)]
    u = np.empty(n * L, dtype=np.uint32)
    assert lib.ya32_sample_uniform(capi.ptr(u), n, L, 5) == 0
    for j, p in enumerate(ps):
        r = u[j * n:(j + 1) * n]
        assert r.max() < p and r.min() >= 0 and abs(r.mean() / p - 0.5) < 0.05
    g = np.empty(n * L, dtype=np.uint32)
    assert lib.ya32_sample_gaussian(capi.ptr(g), n, L, 3.2, 6) == 0
    s0 = np.where(g[:n] > ps[0] // 2, g[:n].astype(np.int64) - ps[0], g[:n].astype(np.int64))
    for j, p in enumerate(ps):
        sj = np.where(g[j * n:(j + 1) * n] > p // 2, g[j * n:(j + 1) * n].astype(np.int64) - p,
                      g[j * n:(j + 1) * n].astype(np.int64))
        assert np.array_equal(sj, s0)
    assert abs(s0.std() - 3.2) < 0.3
    h = np.empty(n * L, dtype=np.uint32)
    assert lib.ya32_sample_hwt(capi.ptr(h), n, L, 64, 7) == 0
    assert np.count_nonzero(h[:n]) == 64 and set(np.unique(h[:n])) <= {0, 1, ps[0] - 1}
    z = np.empty(n * L, dtype=np.uint32)
    assert lib.ya32_sample_zo(capi.ptr(z), n, L, 127, 8) == 0
    assert set(np.unique(z[:n])) <= {0, 1, ps[0] - 1}
    assert lib.ya32_sample_uniform(capi.ptr(u), 1 << 16, 1, 5) == 1      # beyond 2^15
