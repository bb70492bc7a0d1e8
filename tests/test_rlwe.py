"""The reference example's RLWE flow (example.cpp:12-58) on the CPU: the
oracle's restatement (oracle/rlwe_oracle.py over the C oracle) against the
same flow through the reference's own sources (oracle/_ref), bit-exact, and
decryption against the scheme's closed form plain - msg = e1 - u*e0 + sx*e2
(negacyclic products by schoolbook convolution), which fails on a dropped
term, a wrong sign or a swapped operand in the composition."""
import os

import numpy as np
import pytest

from _oracle import REF, REF32, Oracle, RefLib, rlwe_oracle
from yell_amd import capi

R = rlwe_oracle()


def samples(n, L, word, seed):
    """(sx, e0, pk_ax, msg, u, e1, e2) from the host samplers."""
    h = lambda kind, param, s: capi.sample_host(kind, n, L, param, seed + s, word)
    return (h("hwt", min(64, n), 0), h("gaussian", 3.2, 1), h("uniform", None, 2), h("uniform", None, 3),
            h("zo", 0x7F, 4), h("gaussian", 3.2, 5), h("gaussian", 3.2, 6))


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("n,L", [(8, 1), (1024, 2), (8192, 4)])
def test_oracle_flow_matches_reference(word, n, L):
    if not os.path.exists(REF if word == 64 else REF32):
        pytest.skip("oracle/_ref not built")
    orc, ref = Oracle(word), RefLib(word)
    ins = samples(n, L, word, 100 + n)
    got = R.flow(orc, n, L, *ins)
    want = ref.rlwe(n, L, *ins)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("n,L", [(64, 2), (512, 3)])
def test_decryption_closed_form(word, n, L):
    orc = Oracle(word)
    sx, e0, ax, msg, u, e1, e2 = samples(n, L, word, 7 + n)
    plain = R.flow(orc, n, L, sx, e0, ax, msg, u, e1, e2)[4]
    for cm in range(L):
        p = orc.prime(cm)
        s = slice(cm * n, (cm + 1) * n)
        ue0 = orc.schoolbook(u[s], e0[s], n, cm).astype(object)
        sxe2 = orc.schoolbook(sx[s], e2[s], n, cm).astype(object)
        noise = (e1[s].astype(object) - ue0 + sxe2) % p
        assert [int(x) for x in (plain[s].astype(object) - msg[s].astype(object)) % p] == [int(x) for x in noise]
        centred = [x - p if x > p // 2 else x for x in (int(v) for v in noise)]
        assert max(abs(x) for x in centred) < 64 * 13 * 3.2 * 2 + 13 * 3.2


def test_negate_and_lazy_quirks_reached():
    """The flow meets the reference's quirks: forward_lazy leaves words in
    [p, 4p) that go into add_product_of unreduced, and negate maps 0 to p
    (a key word of exactly 0 has probability 1/p, so that case is checked on
    its own)."""
    orc = Oracle(64)
    n, L = 8, 1
    seen_lazy = False
    for seed in range(50):
        sx, e0, ax, msg, u, e1, e2 = samples(n, L, 64, seed)
        lazy = orc.poly_forward_lazy(orc.poly_add(e1, msg, n, L), n, L)
        seen_lazy |= bool(np.any(lazy >= orc.prime(0)))
    assert seen_lazy
    ref_p = [int(x) for x in orc.poly_negate(np.zeros(n, dtype=np.uint64), n, L)]
    assert ref_p == [orc.prime(0)] * n
