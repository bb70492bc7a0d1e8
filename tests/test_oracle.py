"""Pins the CPU oracle (oracle/yell_oracle.c): against Python big integers,
closed forms, the GMP schoolbook, the reference compiled from its own sources
(oracle/_ref, where built) and the reference-produced golden vectors."""
import os

import numpy as np
import pytest

from _oracle import GmpOracle, Oracle, RefLib, rand_residues

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ntt_ref.npz"))
orc = Oracle()
P = [orc.prime(i) for i in range(32)]
M64 = (1 << 64) - 1


def _ref():
    try:
        return RefLib()
    except RuntimeError:
        pytest.skip("oracle/_ref not built here")


def _brv(i, bits):
    return int(format(i, "0%db" % bits)[::-1], 2) if bits else 0


# ---- scalar ops -------------------------------------------------------------
def test_barrett_canonical_and_matches_reference():
    rng = np.random.default_rng(1)
    ref = None
    try:
        ref = RefLib()
    except RuntimeError:
        pass
    for cm in (0, 1, 7, 23, 999):
        p = orc.prime(cm)
        xs = [0, 1, p - 1, p, 2 * p, p * p, (p - 1) ** 2, 16 * p * p - 1, (1 << 128) - 1]
        xs += [int(v) for v in rng.integers(0, 1 << 62, size=200, dtype=np.uint64)]
        xs += [int(a) * int(b) for a, b in rng.integers(0, 4 * p, size=(300, 2), dtype=np.uint64)]
        xs += [(int(a) << 64) | int(b) for a, b in rng.integers(0, M64, size=(300, 2), dtype=np.uint64)]
        for x in xs:
            got = orc.lib.or_barrett_hl(x >> 64, x & M64, cm)
            if x < 16 * p * p:
                assert got == x % p
            if ref is not None:
                assert got == ref.lib.yref_barrett(x >> 64, x & M64, cm)


def test_shoup_ops():
    rng = np.random.default_rng(2)
    for cm in (0, 5, 31):
        p = orc.prime(cm)
        for w in [0, 1, p - 1] + [int(v) for v in rng.integers(0, p, 50, dtype=np.uint64)]:
            ws = orc.lib.or_shoupify(w, cm)
            assert ws == (w << 64) // p
            for x in [0, 4 * p - 1] + [int(v) for v in rng.integers(0, 4 * p, 20, dtype=np.uint64)]:
                assert orc.lib.or_mulmod_shoup(x, w, ws, cm) == x * w % p


def test_add_sub_mul_muladd():
    rng = np.random.default_rng(3)
    for cm in (0, 11):
        p = orc.prime(cm)
        for x, y, z in rng.integers(0, p, size=(300, 3), dtype=np.uint64):
            x, y, z = int(x), int(y), int(z)
            assert orc.lib.or_addmod(x, y, cm) == (x + y) % p
            assert orc.lib.or_submod(x, y, cm) == (x - y) % p
            assert orc.lib.or_mulmod(x, y, cm) == x * y % p
            assert orc.lib.or_muladd(z, x, y, cm) == (x * y + z) % p


# ---- tables -----------------------------------------------------------------
@pytest.mark.parametrize("logn", [3, 5, 10, 14])
@pytest.mark.parametrize("cm", [0, 7])
def test_tables_closed_form(logn, cm):
    n = 1 << logn
    p = orc.prime(cm)
    phi_t, phi_s, iphi_t, iphi_s, sc = orc.tables(n, cm)
    phi = pow(int(orc.lib.or_root(cm)), 1 << (20 - logn), p)
    assert pow(phi, n, p) == p - 1                     # primitive 2N-th root
    idx = list(range(n)) if n <= 1024 else list(range(0, n, 97)) + [n - 1]
    for i in idx:
        j = _brv(i, logn)
        assert int(phi_t[j]) == pow(phi, i, p)
        assert int(iphi_t[j]) == pow(phi, (2 * n - i) % (2 * n), p)
        assert int(phi_s[j]) == (int(phi_t[j]) << 64) // p
        assert int(iphi_s[j]) == (int(iphi_t[j]) << 64) // p
    assert int(sc[0]) * n % p == 1
    assert int(sc[2]) == int(sc[0]) * int(iphi_t[1]) % p
    assert int(sc[1]) == (int(sc[0]) << 64) // p and int(sc[3]) == (int(sc[2]) << 64) // p


@pytest.mark.parametrize("logn", [2, 3, 8, 12, 16])
def test_tables_match_reference(logn):
    ref = _ref()
    n = 1 << logn
    for cm in (0, 3):
        for a, b in zip(orc.tables(n, cm), ref.tables(n, cm)):
            np.testing.assert_array_equal(a, b)


# ---- transforms -------------------------------------------------------------
@pytest.mark.parametrize("logn", [2, 3, 4, 6])
def test_forward_closed_form(logn):
    n = 1 << logn
    rng = np.random.default_rng(logn)
    for cm in (0, 9):
        p = orc.prime(cm)
        a = rng.integers(0, 4 * p, size=n, dtype=np.uint64)
        got = orc.forward(a, n, cm)
        phi = pow(int(orc.lib.or_root(cm)), 1 << (20 - logn), p)
        for j in range(n):
            e = 2 * _brv(j, logn) + 1
            want = sum(int(a[i]) * pow(phi, e * i, p) for i in range(n)) % p
            assert int(got[j]) == want
        lazy = orc.forward_lazy(a, n, cm)
        assert all(int(v) < 4 * p for v in lazy)
        assert [int(v) % p for v in lazy] == [int(v) for v in got]


@pytest.mark.parametrize("logn", [3, 7, 10, 13])
def test_backward_inverts_forward(logn):
    n = 1 << logn
    rng = np.random.default_rng(100 + logn)
    for cm in (0, 2, 17):
        p = orc.prime(cm)
        a = rng.integers(0, p, size=n, dtype=np.uint64)
        np.testing.assert_array_equal(orc.backward(orc.forward(a, n, cm), n, cm), a)


@pytest.mark.parametrize("logn", [2, 3, 4, 9, 12, 14, 18, 19])
def test_transforms_match_reference(logn):
    ref = _ref()
    n = 1 << logn
    rng = np.random.default_rng(200 + logn)
    for cm in ((0, 5, 23) if logn <= 14 else (0,)):
        p = orc.prime(cm)
        x4 = rng.integers(0, 4 * p, size=n, dtype=np.uint64)
        x2 = rng.integers(0, 2 * p, size=n, dtype=np.uint64)
        np.testing.assert_array_equal(orc.forward_lazy(x4, n, cm), ref.forward_lazy(x4, n, cm))
        np.testing.assert_array_equal(orc.forward(x4, n, cm), ref.forward(x4, n, cm))
        np.testing.assert_array_equal(orc.backward(x2, n, cm), ref.backward(x2, n, cm))


def test_golden_vectors():
    for n in (4, 8, 64, 1024):
        for cm in (0, 7):
            k = "n%d_cm%d" % (n, cm)
            np.testing.assert_array_equal(orc.forward_lazy(GOLD["fwd_in_" + k], n, cm),
                                          GOLD["fwd_lazy_out_" + k])
            np.testing.assert_array_equal(orc.forward(GOLD["fwd_in_" + k], n, cm), GOLD["fwd_out_" + k])
            np.testing.assert_array_equal(orc.backward(GOLD["bwd_in_" + k], n, cm), GOLD["bwd_out_" + k])
    for n, L in ((4096, 1), (1024, 3)):
        k = "n%d_L%d" % (n, L)
        np.testing.assert_array_equal(orc.polymul(GOLD["pm_a_" + k], GOLD["pm_b_" + k], n, L),
                                      GOLD["pm_c_" + k])


# ---- poly-multiplication against convolution --------------------------------
def _py_negacyclic(a, b, p):
    n = len(a)
    c = [0] * n
    for i in range(n):
        for j in range(n):
            k = i + j
            if k < n:
                c[k] += int(a[i]) * int(b[j])
            else:
                c[k - n] -= int(a[i]) * int(b[j])
    return [v % p for v in c]


def test_polymul_python_bigint():
    n, L = 32, 3
    rng = np.random.default_rng(5)
    a = rand_residues(rng, n, L, P)
    b = rand_residues(rng, n, L, P)
    c = orc.polymul(a, b, n, L)
    for r in range(L):
        s = slice(r * n, (r + 1) * n)
        assert [int(v) for v in c[s]] == _py_negacyclic(a[s], b[s], P[r])


@pytest.mark.parametrize("n", [256, 4096])
def test_polymul_schoolbook(n):
    rng = np.random.default_rng(n)
    a = rand_residues(rng, n, 1, P)
    b = rand_residues(rng, n, 1, P)
    np.testing.assert_array_equal(orc.polymul(a, b, n, 1), orc.schoolbook(a, b, n, 0))


def test_polymul_gmp():
    try:
        g = GmpOracle()
    except RuntimeError:
        pytest.skip("GMP oracle not built")
    n, L = 1024, 2
    rng = np.random.default_rng(6)
    a = rand_residues(rng, n, L, P)
    b = rand_residues(rng, n, L, P)
    c = orc.polymul(a, b, n, L)
    for r in range(L):
        s = slice(r * n, (r + 1) * n)
        np.testing.assert_array_equal(c[s], g.schoolbook(a[s], b[s], n, P[r]))


def test_pointwise_ops():
    n, L = 64, 4
    rng = np.random.default_rng(7)
    a = rand_residues(rng, n, L, P)
    b = rand_residues(rng, n, L, P)
    acc = rand_residues(rng, n, L, P)
    a[0] = 0
    neg = orc.poly_negate(a, n, L)
    assert int(neg[0]) == P[0]                      # negate keeps p for 0
    for r in range(L):
        for i in range(n):
            k = r * n + i
            x, y, z, p = int(a[k]), int(b[k]), int(acc[k]), P[r]
            assert int(orc.poly_mul(a, b, n, L)[k]) == x * y % p
            assert int(orc.poly_add(a, b, n, L)[k]) == (x + y) % p
            assert int(orc.poly_sub(a, b, n, L)[k]) == (x - y) % p
            assert int(orc.poly_add_product_of(acc, a, b, n, L)[k]) == (z + x * y) % p
            if i > 8:
                break
