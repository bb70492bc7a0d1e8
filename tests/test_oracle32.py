"""The 32-bit moduli oracle (liboracle32.so, the restatement compiled with
32-bit words) against the reference's own 32-bit build (golden vectors and,
where built, oracle/_ref/libyellref32.so) and against the mathematics."""
import os

import numpy as np
import pytest

from _oracle import REF32, Oracle, RefLib

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ntt32_ref.npz"))
O = Oracle(32)
HAVE_REF = os.path.exists(REF32)
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libyellref32.so not built")


def brv(j, logn):
    return int(format(j, "0%db" % logn)[::-1], 2) if logn else 0


def test_constants():
    assert O.lib.or_num_moduli() == 291
    assert O.prime(0) == 1073479681 and O.prime(290) == 873005057


@pytest.mark.parametrize("cm", [0, 3, 290])
def test_scalar_ops_closed_forms(cm):
    rng = np.random.default_rng(cm)
    p = O.prime(cm)
    for _ in range(2000):
        x, y, z = (int(v) for v in rng.integers(0, p, size=3))
        assert O.lib.or_mulmod(x, y, cm) == x * y % p
        assert O.lib.or_muladd(z, x, y, cm) == (x * y + z) % p
        assert O.lib.or_addmod(x, y, cm) == (x + y) % p
        assert O.lib.or_submod(x, y, cm) == (x - y) % p
        ws = O.lib.or_shoupify(y, cm)
        assert ws == (y << 32) // p
        x4 = int(rng.integers(0, 4 * p))
        assert O.lib.or_mulmod_shoup(x4, y, ws, cm) == x4 * y % p
    for hi in (0, 1, p - 1, (p * p) >> 32):
        for lo in (0, 1, 0xFFFFFFFF):
            v = (hi << 32) | lo
            if v < p * p + p:
                assert O.lib.or_barrett_hl(hi, lo, cm) == v % p


@needs_ref
@pytest.mark.parametrize("cm", [0, 290])
def test_scalar_ops_match_reference(cm):
    ref = RefLib(32)
    rng = np.random.default_rng(100 + cm)
    p = O.prime(cm)
    for _ in range(3000):
        x, y, z = (int(v) for v in rng.integers(0, p, size=3))
        assert O.lib.or_mulmod(x, y, cm) == ref.lib.yref_mulmod(x, y, cm)
        assert O.lib.or_muladd(z, x, y, cm) == ref.lib.yref_muladd(z, x, y, cm)
        ws = O.lib.or_shoupify(y, cm)
        assert ws == ref.lib.yref_shoupify(y, cm)
        x4 = int(rng.integers(0, 4 * p))
        assert O.lib.or_mulmod_shoup(x4, y, ws, cm) == ref.lib.yref_mulmod_shoup(x4, y, ws, cm)
        hi, lo = int(rng.integers(0, 1 << 32)), int(rng.integers(0, 1 << 32))
        assert O.lib.or_barrett_hl(hi, lo, cm) == ref.lib.yref_barrett(hi, lo, cm)


@pytest.mark.parametrize("logn", [3, 5, 10, 15])
@pytest.mark.parametrize("cm", [0, 290])
def test_tables_closed_form(logn, cm):
    n = 1 << logn
    p = O.prime(cm)
    phi, phis, iphi, iphis, sc = O.tables(n, cm)
    g = int(phi[brv(1, logn)])                           # phi^1 sits at brv(1)
    assert pow(g, n, p) == p - 1
    for i in (0, 1, 2, n - 1):
        assert int(phi[brv(i, logn)]) == pow(g, i, p)
        assert int(iphi[brv(i, logn)]) == pow(g, -i, p)
    assert all(int(phis[i]) == (int(phi[i]) << 32) // p for i in range(0, n, max(1, n // 64)))
    assert int(sc[0]) * n % p == 1
    assert int(sc[2]) == int(sc[0]) * int(iphi[1]) % p


@needs_ref
@pytest.mark.parametrize("logn", list(range(3, 16)))
def test_tables_match_reference(logn):
    ref = RefLib(32)
    for cm in (0, 290):
        for a, b in zip(O.tables(1 << logn, cm), ref.tables(1 << logn, cm)):
            np.testing.assert_array_equal(a, b)


def test_forward_closed_form_small():
    for logn in (2, 3, 4, 5):
        n = 1 << logn
        cm = 7
        p = O.prime(cm)
        rng = np.random.default_rng(logn)
        a = rng.integers(0, p, size=n, dtype=np.uint32)
        phi = O.tables(n, cm)[0]
        g = int(phi[brv(1, logn)])
        got = O.forward(a, n, cm)
        for j in range(n):
            e = 2 * brv(j, logn) + 1
            assert int(got[j]) == sum(int(a[i]) * pow(g, e * i, p) for i in range(n)) % p


@pytest.mark.parametrize("key", sorted({k[len("fwd_in_"):] for k in GOLD.files if k.startswith("fwd_in_")}))
def test_transforms_match_golden(key):
    n, cm = (int(x[1:] if x[0] == "n" else x[2:]) for x in key.split("_"))
    np.testing.assert_array_equal(O.forward_lazy(GOLD["fwd_in_" + key], n, cm), GOLD["fwd_lazy_out_" + key])
    np.testing.assert_array_equal(O.forward(GOLD["fwd_in_" + key], n, cm), GOLD["fwd_out_" + key])
    np.testing.assert_array_equal(O.backward(GOLD["bwd_in_" + key], n, cm), GOLD["bwd_out_" + key])


@pytest.mark.parametrize("key", sorted({k[5:] for k in GOLD.files if k.startswith("pm_c_")}))
def test_polymul_matches_golden(key):
    n, L = (int(x[1:]) for x in key.split("_"))
    got = O.polymul(GOLD["pm_a_" + key], GOLD["pm_b_" + key], n, L)
    np.testing.assert_array_equal(got, GOLD["pm_c_" + key])


def test_polymul_schoolbook_and_roundtrip():
    n, cm = 256, 5
    p = O.prime(cm)
    rng = np.random.default_rng(3)
    a = rng.integers(0, p, size=n, dtype=np.uint32)
    b = rng.integers(0, p, size=n, dtype=np.uint32)
    want = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                want[k] += int(a[i]) * int(b[j])
            else:
                want[k - n] -= int(a[i]) * int(b[j])
    want = np.array([v % p for v in want], dtype=np.uint32)
    np.testing.assert_array_equal(O.schoolbook(a, b, n, cm), want)
    # polymul on residue cm: use a poly whose residue index equals cm
    L = cm + 1
    A = np.zeros(n * L, dtype=np.uint32)
    B = np.zeros(n * L, dtype=np.uint32)
    A[cm * n:], B[cm * n:] = a, b
    np.testing.assert_array_equal(O.polymul(A, B, n, L)[cm * n:], want)
    x = rng.integers(0, p, size=n, dtype=np.uint32)
    np.testing.assert_array_equal(O.backward(O.forward(x, n, cm), n, cm), x)


@needs_ref
@pytest.mark.parametrize("logn", [3, 6, 9, 12, 15])
def test_transforms_match_reference_library(logn):
    ref = RefLib(32)
    n = 1 << logn
    rng = np.random.default_rng(logn)
    for cm in (1, 200):
        p = O.prime(cm)
        x4 = rng.integers(0, 4 * p, size=n, dtype=np.uint32)
        x2 = rng.integers(0, 2 * p, size=n, dtype=np.uint32)
        np.testing.assert_array_equal(O.forward_lazy(x4, n, cm), ref.forward_lazy(x4, n, cm))
        np.testing.assert_array_equal(O.backward(x2, n, cm), ref.backward(x2, n, cm))
