#!/usr/bin/env python3
"""Write tests/golden/ntt_ref.npz: input/output vectors of the reference.

The outputs come from the reference's own lib/ntt.cpp, lib/ops.cpp,
lib/math.cpp and lib/params.cpp compiled by oracle/Makefile into
oracle/_ref/libyellref.so (ntt<N>::forward_lazy / forward / backward and the
poly-multiplication call sequence).  Inputs are seeded numpy draws: forward
inputs in [0, 4p), backward inputs in [0, 2p), poly-mult operands in [0, p).
Run in the container that holds the reference checkout; commit the .npz.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _oracle import Oracle, RefLib  # noqa: E402

SEED = 20261203


def main():
    ref = RefLib()
    rng = np.random.default_rng(SEED)
    out = {}
    orc = Oracle()
    for n in (4, 8, 64, 1024):
        for cm in (0, 7):
            q = orc.prime(cm)
            x4 = rng.integers(0, 4 * q, size=n, dtype=np.uint64)
            x2 = rng.integers(0, 2 * q, size=n, dtype=np.uint64)
            out["fwd_in_n%d_cm%d" % (n, cm)] = x4
            out["fwd_lazy_out_n%d_cm%d" % (n, cm)] = ref.forward_lazy(x4, n, cm)
            out["fwd_out_n%d_cm%d" % (n, cm)] = ref.forward(x4, n, cm)
            out["bwd_in_n%d_cm%d" % (n, cm)] = x2
            out["bwd_out_n%d_cm%d" % (n, cm)] = ref.backward(x2, n, cm)
    for n, L in ((4096, 1), (1024, 3)):
        a = np.concatenate([rng.integers(0, orc.prime(r), size=n, dtype=np.uint64) for r in range(L)])
        b = np.concatenate([rng.integers(0, orc.prime(r), size=n, dtype=np.uint64) for r in range(L)])
        out["pm_a_n%d_L%d" % (n, L)] = a
        out["pm_b_n%d_L%d" % (n, L)] = b
        out["pm_c_n%d_L%d" % (n, L)] = ref.polymul(a, b, n, L)
    np.savez_compressed(os.path.join(HERE, "ntt_ref.npz"), **out)


if __name__ == "__main__":
    main()
