#!/usr/bin/env python3
"""Write tests/golden/crt_ref.npz: CRT-lift vectors of the reference.

The outputs come from the reference's own lib/gmp.cpp (gmp::init_table) and
poly<N>::poly2mpz (include/yell/gmp.hpp:41-70), compiled with lib/mem.cpp,
lib/ntt.cpp, lib/ops.cpp, lib/math.cpp and lib/params.cpp by oracle/Makefile
into oracle/_ref/libyellref_gmp.so.  Each lifted coefficient is stored as
little-endian 64-bit limbs.  Inputs are seeded numpy draws: canonical
residues (< P[j]), arbitrary 64-bit words, and edge words (0, P[j]-1, P[j],
2^64-1).  Run in the container that holds the reference checkout; commit the
.npz.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _oracle import Oracle, RefGmpLib, crt_oracle  # noqa: E402

SEED = 20261204
# (log2 degree, nmoduli, kind): kind 0 canonical, 1 any 64-bit word, 2 edge words
CASES = [(3, 1, 0), (4, 2, 1), (5, 3, 2), (8, 8, 0), (6, 8, 1), (5, 12, 2), (4, 24, 0),
         (4, 32, 1), (3, 33, 0), (3, 40, 1), (4, 8, 2)]
CONSTANT_L = (1, 2, 3, 8, 24, 32, 33, 64)


def main():
    ref = RefGmpLib()
    orc = Oracle()
    crt = crt_oracle()
    rng = np.random.default_rng(SEED)
    out = {}
    for logn, L, kind in CASES:
        n = 1 << logn
        primes = [orc.prime(j) for j in range(L)]
        if kind == 0:
            res = np.concatenate([rng.integers(0, p, size=n, dtype=np.uint64) for p in primes])
        elif kind == 1:
            res = rng.integers(0, 2**64 - 1, size=n * L, dtype=np.uint64, endpoint=True)
        else:
            res = np.empty(n * L, dtype=np.uint64)
            for j, p in enumerate(primes):
                edges = np.array([0, p - 1, p, 2**64 - 1, 1, 2 * p, 4 * p - 1, p + 1], dtype=np.uint64)
                res[j * n:(j + 1) * n] = edges[rng.integers(0, len(edges), size=n)]
        limbs = crt.limbs_for(primes)
        key = "n%d_L%d_k%d" % (n, L, kind)
        out["in_" + key] = res
        out["out_" + key] = ref.poly2mpz(res, n, L, limbs)
    for L in CONSTANT_L:
        primes = [orc.prime(j) for j in range(L)]
        limbs = crt.limbs_for(primes)
        q, m, lift, shift, bits = ref.constants(L, limbs)
        out["q_L%d" % L] = q
        out["m_L%d" % L] = m
        out["lift_L%d" % L] = lift
        out["shift_bits_L%d" % L] = np.array([shift, bits], dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "crt_ref.npz"), **out)


if __name__ == "__main__":
    main()
