"""GPU CRT lift (ya_poly_crt_lift, the device form of poly2mpz) against the
reference's own vectors, the Python restatement and the CRT itself."""
import ctypes
import os

import numpy as np
import pytest

from _oracle import Oracle, crt_oracle
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "crt_ref.npz"))
CRT = crt_oracle()
ORC = Oracle()


def primes(L):
    return [ORC.prime(j) for j in range(L)]


def limbs(L):
    w = ctypes.c_size_t()
    capi.call("ya_crt_limbs", L, ctypes.byref(w))
    return w.value


def gpu_lift(res, n, L):
    device()
    w = limbs(L)
    din = capi.DeviceBuffer.from_host(np.ascontiguousarray(res, dtype=np.uint64))
    dout = capi.DeviceBuffer(n * w)
    capi.call("ya_poly_crt_lift", dout.ptr, din.ptr, n, L, None)
    out = dout.download().reshape(n, w)
    din.close()
    dout.close()
    return out


@pytest.mark.parametrize("key", [k[3:] for k in GOLD.files if k.startswith("in_")])
def test_lift_matches_reference_vectors(key):
    n, L, kind = (int(x[1:]) for x in key.split("_"))
    say("crt golden", key)
    assert np.array_equal(gpu_lift(GOLD["in_" + key], n, L), GOLD["out_" + key])


def _check_against_oracle(res, n, L, sample=None):
    got = gpu_lift(res, n, L)
    idx = np.arange(n) if sample is None else sample
    sub = np.asarray(res, dtype=np.uint64).reshape(L, n)[:, idx]
    want = CRT.to_limbs(CRT.lift_values(sub, len(idx), primes(L)), got.shape[1])
    assert np.array_equal(got[idx], want)
    return got


@pytest.mark.parametrize("n,L", [(16384, 8), (4096, 1), (8192, 3), (2048, 12), (1024, 16), (512, 32)])
def test_lift_register_path(n, L):
    say("crt register path", n, L)
    rng = np.random.default_rng(n + L)
    res = np.concatenate([rng.integers(0, p, size=n, dtype=np.uint64) for p in primes(L)])
    got = _check_against_oracle(res, n, L)
    # canonical residues: the lift reduces back to them
    vals = CRT.from_limbs(got[:64])
    for j, p in enumerate(primes(L)):
        assert [v % int(p) for v in vals] == [int(x) for x in res[j * n:j * n + 64]]


def test_lift_c4_shape():
    """N = 2^16, L = 24: every 61st coefficient against the restatement."""
    n, L = 1 << 16, 24
    say("crt 2^16 x 24 shape", n, L)
    rng = np.random.default_rng(24)
    res = np.concatenate([rng.integers(0, p, size=n, dtype=np.uint64) for p in primes(L)])
    _check_against_oracle(res, n, L, sample=np.arange(0, n, 61))


@pytest.mark.parametrize("n,L", [(256, 33), (64, 64), (16, 200), (8, 1000), (9000, 40)])
def test_lift_scratch_path(n, L):
    say("crt scratch path", n, L)
    rng = np.random.default_rng(n * 7 + L)
    res = rng.integers(0, 2**64 - 1, size=n * L, dtype=np.uint64, endpoint=True)
    sample = None if n <= 256 else np.concatenate([np.arange(0, 64), np.arange(8190, 8200), np.arange(n - 16, n)])
    _check_against_oracle(res, n, L, sample=sample)


@pytest.mark.parametrize("n", [1, 7, 1000, 8193])
def test_lift_any_coefficient_count(n):
    say("crt count", n)
    L = 4
    rng = np.random.default_rng(n)
    res = rng.integers(0, 2**64 - 1, size=n * L, dtype=np.uint64, endpoint=True)
    _check_against_oracle(res, n, L)


def test_lift_extremes():
    """All-zero, all (P[j] - 1) (the lift is Q - 1) and all 2^64 - 1 residues."""
    n, L = 256, 8
    ps = primes(L)
    q = 1
    for p in ps:
        q *= int(p)
    z = gpu_lift(np.zeros(n * L, dtype=np.uint64), n, L)
    assert not z.any()
    top = np.concatenate([np.full(n, p - 1, dtype=np.uint64) for p in ps])
    assert CRT.from_limbs(gpu_lift(top, n, L)) == [q - 1] * n
    ones = np.full(n * L, 2**64 - 1, dtype=np.uint64)
    _check_against_oracle(ones, n, L)


def test_lift_rejects():
    device()
    a = capi.DeviceBuffer(64)
    lib = capi.load()
    assert lib.ya_poly_crt_lift(a.ptr, a.ptr, 8, 2, None) == 1           # overlap
    assert lib.ya_poly_crt_lift(a.ptr, a.ptr, 0, 2, None) == 1           # empty
    assert lib.ya_poly_crt_lift(a.ptr, a.ptr, 8, 1001, None) == 1        # too many primes
    a.close()
