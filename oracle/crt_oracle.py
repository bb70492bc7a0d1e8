"""CRT lift restated with Python integers -- TEST INFRASTRUCTURE ONLY.

Only tests/, the package's smoke check and the bench's CPU baseline may use
this module, and only as a checker.

Follows the reference step by step:
  constants  gmp::gmp_precomputed::gmp_precomputed (lib/gmp.cpp:6-51):
             Q = prod P[j]; bits = bitsize(Q);
             shift = bits + kModulusRepresentationBitsize (64) + (T)log2(L) + 1;
             M = floor(2^shift / Q);
             lift_j = (Q / P[j]) * ((Q / P[j])^-1 mod P[j]).
  lift       gmp::poly2mpz (include/yell/gmp.hpp:41-70), per coefficient:
             S = sum_j lift_j * r_j;  tmp = floor(S * M / 2^shift);
             S -= tmp * Q;  S -= Q if S >= Q.
Pinned (tests/test_crt.py) by: the reference's own gmp.cpp / poly2mpz
compiled in oracle/_ref (constants and lifted values, tests/golden/crt_ref.npz),
and by the Chinese remainder theorem itself -- the result x satisfies
0 <= x < Q and x = r_j (mod P[j]) for every j when r_j < P[j], and
x = sum_j lift_j r_j mod Q for any 64-bit residues.
"""
import numpy as np


def constants(primes):
    """(Q, bits, shift, M, lifts) for the primes P[0..L)."""
    q = 1
    for p in primes:
        q *= int(p)
    bits = q.bit_length()
    L = len(primes)
    shift = bits + 64 + (L.bit_length() - 1) + 1
    m = (1 << shift) // q
    lifts = []
    for p in primes:
        p = int(p)
        quo = q // p
        lifts.append(quo * pow(quo % p, -1, p))
    return q, bits, shift, m, lifts


def lift_values(residues, n, primes):
    """residues: [L][n] words (flat or 2-D) -> list of n Python ints."""
    q, _, shift, m, lifts = constants(primes)
    r = np.asarray(residues, dtype=np.uint64).reshape(len(primes), n)
    cols = [[int(v) for v in r[j]] for j in range(len(primes))]
    out = []
    for i in range(n):
        s = 0
        for j, lj in enumerate(lifts):
            v = cols[j][i]
            if v:
                s += lj * v
        tmp = (s * m) >> shift
        s -= tmp * q
        if s >= q:
            s -= q
        out.append(s)
    return out


def limbs_for(primes):
    q, bits, _, _, _ = constants(primes)
    return (bits + 63) // 64


def to_limbs(values, limbs):
    """List of non-negative ints -> uint64 [len][limbs], little-endian limbs."""
    out = np.zeros((len(values), limbs), dtype=np.uint64)
    mask = (1 << 64) - 1
    for i, v in enumerate(values):
        for k in range(limbs):
            out[i, k] = (v >> (64 * k)) & mask
        if v >> (64 * limbs):
            raise ValueError("value does not fit in %d limbs" % limbs)
    return out


def from_limbs(arr):
    """uint64 [n][limbs] -> list of ints."""
    arr = np.asarray(arr, dtype=np.uint64)
    vals = []
    for row in arr:
        v = 0
        for k in range(len(row) - 1, -1, -1):
            v = (v << 64) | int(row[k])
        vals.append(v)
    return vals
