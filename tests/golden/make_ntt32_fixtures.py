#!/usr/bin/env python3
"""Write tests/golden/ntt32_ref.npz: vectors of the reference's 32-bit moduli
build (YELL_USE_32BITS_MODULI).

The outputs come from the reference's own lib/ntt.cpp, lib/ops.cpp,
lib/math.cpp and lib/params.cpp compiled with -DYELL_USE_32BITS_MODULI by
oracle/Makefile into oracle/_ref/libyellref32.so (ntt<N>::forward_lazy /
forward / backward and the poly-multiplication call sequence, uint32 words).
Inputs are seeded numpy draws: forward inputs in [0, 4p), backward inputs in
[0, 2p), poly-mult operands in [0, p).  Run in the container that holds the
reference checkout; commit the .npz.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _oracle import Oracle, RefLib  # noqa: E402

SEED = 20261205


def main():
    ref = RefLib(32)
    orc = Oracle(32)
    rng = np.random.default_rng(SEED)
    out = {}
    for n in (8, 64, 1024, 4096):
        for cm in (0, 290):
            q = orc.prime(cm)
            x4 = rng.integers(0, 4 * q, size=n, dtype=np.uint32)
            x2 = rng.integers(0, 2 * q, size=n, dtype=np.uint32)
            key = "n%d_cm%d" % (n, cm)
            out["fwd_in_" + key] = x4
            out["fwd_lazy_out_" + key] = ref.forward_lazy(x4, n, cm)
            out["fwd_out_" + key] = ref.forward(x4, n, cm)
            out["bwd_in_" + key] = x2
            out["bwd_out_" + key] = ref.backward(x2, n, cm)
    for n, L in ((4096, 1), (8192, 2), (1024, 3)):
        a = np.concatenate([rng.integers(0, orc.prime(r), size=n, dtype=np.uint32) for r in range(L)])
        b = np.concatenate([rng.integers(0, orc.prime(r), size=n, dtype=np.uint32) for r in range(L)])
        out["pm_a_n%d_L%d" % (n, L)] = a
        out["pm_b_n%d_L%d" % (n, L)] = b
        out["pm_c_n%d_L%d" % (n, L)] = ref.polymul(a, b, n, L)
    np.savez_compressed(os.path.join(HERE, "ntt32_ref.npz"), **out)


if __name__ == "__main__":
    main()
