"""The generated constant tables against the reference's (tests/golden/params_ref.json)
and against what the mathematics fixes (primality, root order, Barrett constant)."""
import hashlib
import json
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(os.path.dirname(HERE), "yell_amd", "csrc", "params_table.inc")


def _table(name):
    s = open(TABLE).read()
    m = re.search(r"#define " + name + r" \{(.*?)\}", s, re.S)
    return [int(x.strip().rstrip("ULL"), 16) for x in m.group(1).replace("\\", "").split(",")
            if x.strip()]


P, PN, ROOT = _table("YA_TABLE_P"), _table("YA_TABLE_PN"), _table("YA_TABLE_ROOT")
INVDEG = _table("YA_TABLE_INVDEG")
FIX = json.load(open(os.path.join(HERE, "golden", "params_ref.json")))


def _digest(v):
    return hashlib.sha256(",".join(str(x) for x in v).encode()).hexdigest()


@pytest.mark.parametrize("ours,name", [(P, "P"), (PN, "Pn"), (ROOT, "primitive_roots"),
                                       (INVDEG, "invkMaxPolyDegree")])
def test_matches_reference_tables(ours, name):
    assert len(ours) == FIX["count"] == 1000
    assert [str(x) for x in ours[:64]] == FIX[name + "_first64"]
    assert _digest(ours) == FIX[name + "_sha256"]


def _is_prime(n):
    d, s = n - 1, 0
    while d % 2 == 0:
        d, s = d // 2, s + 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_prime_form_and_properties():
    for i, (p, pn, r) in enumerate(zip(P[:64], PN[:64], ROOT[:64])):
        assert _is_prime(p)
        assert (p - 1) % (1 << 21) == 0 and (1 << 61) < p < (1 << 62)
        assert pn == ((1 << 128) // p) & ((1 << 64) - 1)
        # primitive 2^21-th root of unity: r^(2^20) == -1
        assert pow(r, 1 << 20, p) == p - 1
    assert all(a > b for a, b in zip(P, P[1:]))   # descending
    # consecutive: no prime of the form skipped between P[0] and P[7]
    k = [((1 << 62) + 1 - p) >> 21 for p in P[:8]]
    for kk in range(1, k[-1]):
        if kk not in k:
            assert not _is_prime((1 << 62) - kk * (1 << 21) + 1)


# ---- 32-bit moduli build (include/yell/params_32.hpp) -------------------------
TABLE32 = os.path.join(os.path.dirname(HERE), "yell_amd", "csrc", "params32_table.inc")


def _table32(name):
    s = open(TABLE32).read()
    m = re.search(r"#define " + name + r" \{(.*?)\}", s, re.S)
    return [int(x.strip().rstrip("u"), 16) for x in m.group(1).replace("\\", "").split(",") if x.strip()]


P32, PN32, ROOT32 = _table32("YA32_TABLE_P"), _table32("YA32_TABLE_PN"), _table32("YA32_TABLE_ROOT")
INVDEG32 = _table32("YA32_TABLE_INVDEG")
FIX32 = json.load(open(os.path.join(HERE, "golden", "params32_ref.json")))


@pytest.mark.parametrize("name", ["P", "Pn", "primitive_roots", "invkMaxPolyDegree"])
def test_32bit_tables_match_reference(name):
    ours = {"P": P32, "Pn": PN32, "primitive_roots": ROOT32, "invkMaxPolyDegree": INVDEG32}[name]
    assert len(ours) == FIX32["count"] == 291
    assert [str(x) for x in ours[:64]] == FIX32[name + "_first64"]
    assert _digest(ours) == FIX32[name + "_sha256"]


def test_32bit_prime_form_and_properties():
    for p, pn, r in zip(P32, PN32, ROOT32):
        assert _is_prime(p)
        assert (p - 1) % (1 << 16) == 0 and (1 << 29) < p < (1 << 30)
        assert pn == ((1 << 64) // p) & 0xffffffff
        assert pow(r, 1 << 15, p) == p - 1          # primitive 2^16-th root
    assert all(a > b for a, b in zip(P32, P32[1:]))
    # invkMaxPolyDegree (params_32.hpp:33) is 2^15 inverted mod each prime
    inv = [pow(1 << 15, -1, p) for p in P32]
    assert _digest(inv) == FIX32["invkMaxPolyDegree_sha256"]
