"""CRT lift (poly2mpz) on the CPU side: the Python restatement in
oracle/crt_oracle.py against the reference's own gmp.cpp / poly2mpz (golden
vectors and, where it is built, oracle/_ref/libyellref_gmp.so), against the
Chinese remainder theorem itself, and the C ABI's host constants."""
import ctypes
import os

import numpy as np
import pytest

from _oracle import REF_GMP, Oracle, RefGmpLib, crt_oracle
from yell_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "crt_ref.npz"))
CRT = crt_oracle()
ORC = Oracle()


def primes(L):
    return [ORC.prime(j) for j in range(L)]


def golden_cases():
    for k in GOLD.files:
        if k.startswith("in_"):
            key = k[3:]
            n, L, kind = (int(x[1:]) for x in key.split("_"))
            yield key, n, L, kind


@pytest.mark.parametrize("L", [1, 2, 3, 8, 24, 32, 33, 64])
def test_oracle_constants_match_reference(L):
    q, bits, shift, m, lifts = CRT.constants(primes(L))
    limbs = CRT.limbs_for(primes(L))
    assert CRT.from_limbs([GOLD["q_L%d" % L]])[0] == q
    assert CRT.from_limbs([GOLD["m_L%d" % L]])[0] == m
    assert CRT.from_limbs(GOLD["lift_L%d" % L]) == lifts
    assert list(GOLD["shift_bits_L%d" % L]) == [shift, bits]
    assert GOLD["lift_L%d" % L].shape == (L, limbs)


@pytest.mark.parametrize("L", [1, 2, 3, 8, 24, 33])
def test_constants_closed_form(L):
    """lift_j = 1 mod P[j], 0 mod P[i] (i != j), < Q; M = floor(2^shift / Q)."""
    ps = primes(L)
    q, bits, shift, m, lifts = CRT.constants(ps)
    assert bits == q.bit_length() and q.bit_length() <= 62 * L
    assert m * q <= (1 << shift) < (m + 1) * q
    for j, lj in enumerate(lifts):
        assert 0 < lj < q
        for i, p in enumerate(ps):
            assert lj % int(p) == (1 if i == j else 0)


@pytest.mark.parametrize("case", list(golden_cases()), ids=lambda c: c[0])
def test_oracle_lift_matches_reference_vectors(case):
    key, n, L, kind = case
    vals = CRT.lift_values(GOLD["in_" + key], n, primes(L))
    assert np.array_equal(CRT.to_limbs(vals, GOLD["out_" + key].shape[1]), GOLD["out_" + key])


@pytest.mark.parametrize("case", list(golden_cases()), ids=lambda c: c[0])
def test_reference_vectors_are_crt(case):
    """The reference's outputs are sum_j lift_j r_j mod Q, in [0, Q), and for
    canonical residues the unique x with x = r_j (mod P[j])."""
    key, n, L, kind = case
    ps = primes(L)
    q, _, _, _, lifts = CRT.constants(ps)
    res = GOLD["in_" + key].reshape(L, n)
    vals = CRT.from_limbs(GOLD["out_" + key])
    for i in range(n):
        x = vals[i]
        assert 0 <= x < q
        assert x == sum(lifts[j] * int(res[j, i]) for j in range(L)) % q
        for j in range(L):
            assert x % int(ps[j]) == int(res[j, i]) % int(ps[j])


def _lift_mutant(res, n, ps, mutation):
    q, _, shift, m, lifts = CRT.constants(ps)
    if mutation == "lift_order":
        lifts = lifts[::-1]
    r = np.asarray(res, dtype=np.uint64).reshape(len(ps), n)
    out = []
    for i in range(n):
        s = sum(lj * int(r[j, i]) for j, lj in enumerate(lifts))
        tmp = (s * m) >> shift
        s -= tmp * q
        if mutation != "no_correction" and s >= q:
            s -= q
        out.append(s)
    return out


@pytest.mark.parametrize("mutation", ["no_correction", "lift_order"])
def test_planted_mistakes_fail_the_vectors(mutation):
    """Mistakes planted in a copy of the restatement (the final conditional
    subtraction dropped; the lifting integers applied to the wrong primes)
    disagree with at least one reference vector."""
    failures = 0
    for key, n, L, kind in golden_cases():
        if L == 1 and mutation == "lift_order":
            continue
        got = CRT.to_limbs(_lift_mutant(GOLD["in_" + key], n, primes(L), mutation), GOLD["out_" + key].shape[1]) \
            if mutation == "lift_order" else None
        if mutation == "no_correction":
            vals = _lift_mutant(GOLD["in_" + key], n, primes(L), mutation)
            failures += vals != CRT.from_limbs(GOLD["out_" + key])
        else:
            failures += not np.array_equal(got, GOLD["out_" + key])
    assert failures > 0


@pytest.mark.skipif(not os.path.exists(REF_GMP), reason="oracle/_ref/libyellref_gmp.so not built")
@pytest.mark.parametrize("logn,L", [(10, 8), (9, 24), (6, 33), (3, 100)])
def test_oracle_lift_matches_reference_library(logn, L):
    ref = RefGmpLib()
    n = 1 << logn
    rng = np.random.default_rng(logn * 1000 + L)
    res = rng.integers(0, 2**64 - 1, size=n * L, dtype=np.uint64, endpoint=True)
    limbs = CRT.limbs_for(primes(L))
    got = CRT.to_limbs(CRT.lift_values(res, n, primes(L)), limbs)
    assert np.array_equal(got, ref.poly2mpz(res, n, L, limbs))


@pytest.mark.parametrize("L", [1, 2, 3, 8, 24, 32, 33, 64, 1000])
def test_capi_constants(L):
    lib = capi.load()
    w = ctypes.c_size_t()
    assert lib.ya_crt_limbs(L, ctypes.byref(w)) == 0
    ps = primes(L)
    assert w.value == CRT.limbs_for(ps)
    q = np.empty(w.value, dtype=np.uint64)
    m = np.empty(2, dtype=np.uint64)
    lift = np.empty((L, w.value), dtype=np.uint64)
    shift = ctypes.c_size_t()
    assert lib.ya_crt_constants(L, w.value, capi.ptr(q), capi.ptr(m), capi.ptr(lift), ctypes.byref(shift)) == 0
    if "q_L%d" % L in GOLD.files:
        assert np.array_equal(q, GOLD["q_L%d" % L])
        assert np.array_equal(m, GOLD["m_L%d" % L])
        assert np.array_equal(lift, GOLD["lift_L%d" % L])
        assert shift.value == int(GOLD["shift_bits_L%d" % L][0])
    qq, bits, sh, mm, lifts = CRT.constants(ps)
    assert CRT.from_limbs([q])[0] == qq and CRT.from_limbs([m])[0] == mm and sh == shift.value
    assert CRT.from_limbs(lift[:4]) == lifts[:4]


def test_capi_crt_rejects():
    lib = capi.load()
    w = ctypes.c_size_t()
    assert lib.ya_crt_limbs(0, ctypes.byref(w)) == 1
    assert lib.ya_crt_limbs(1001, ctypes.byref(w)) == 1
    assert lib.ya_crt_limbs(8, None) == 1
    assert lib.ya_crt_constants(8, 3, None, None, None, None) == 1   # wrong limb count
    assert lib.ya_poly_crt_lift(None, None, 8, 8, None) == 1
