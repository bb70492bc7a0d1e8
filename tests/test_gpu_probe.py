"""VALU ceilings (ya_probe_valu): every kind reports a rate between zero and
the device's lane-instruction rate (256 CUs x 4 SIMDs x 16 lanes at most
~2.5 GHz, ~41 T/s); a butterfly, many instructions long, retires slower than
one multiply-add; an unknown kind is refused."""
import ctypes

import pytest

from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
LANE_OPS_PER_S = 256 * 4 * 16 * 2.5e9


def _probe(kind):
    v = ctypes.c_double()
    capi.call("ya_probe_valu", kind, ctypes.byref(v))
    return v.value


def test_probe_rates_are_physical():
    device()
    say("VALU probes")
    rates = {k: _probe(k) for k in range(4)}
    for k, r in rates.items():
        assert 0 < r < LANE_OPS_PER_S, (k, r)
    assert rates[2] < rates[0] / 8 and rates[3] < rates[0] / 2, rates
    assert rates[3] > rates[2], rates  # the 32-bit butterfly is the shorter sequence
    assert capi.load().ya_probe_valu(4, ctypes.byref(ctypes.c_double())) != 0
