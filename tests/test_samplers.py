"""Host samplers (ya_sample_*, ya32_sample_*) against the sampler oracle
(numpy's Philox + the reference's mappings, details/poly.hpp:200-330): bit-exact
words for the same seed, the Gaussian table against exact values, and the
closed-form properties that pin each mapping."""
import numpy as np
import pytest

from _oracle import Oracle, sampler_oracle
from yell_amd import capi

S = sampler_oracle()
O32 = Oracle(32)


def primes(L, word):
    return [int(p) for p in capi.primes(L)] if word == 64 else [O32.prime(j) for j in range(L)]


def test_philox_known_answer():
    """numpy's Philox4x64-10 at counter 0, key 0 and at all-ones against the
    Random123 known-answer vectors; the oracle's block numbering on top."""
    g = np.random.Philox(key=0, counter=(1 << 256) - 1)
    assert [int(x) for x in g.random_raw(4)] == [0x16554d9eca36314c, 0xdb20fe9d672d0fdc,
                                                 0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    g = np.random.Philox(key=(1 << 128) - 1, counter=(1 << 256) - 2)
    assert [int(x) for x in g.random_raw(4)] == [0x87b092c3013fe90b, 0x438c3c67be8d0224,
                                                 0x9cc7d7c69cd777b6, 0xa09caebf594f0ba0]
    assert [int(x) for x in S.stream_words(0, 4)] == [0x16554d9eca36314c, 0xdb20fe9d672d0fdc,
                                                     0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    w = S.stream_words(7, 12)
    np.testing.assert_array_equal(S.stream_words(7, 4, start_block=2), w[8:12])


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("n,L,seed", [(8, 1, 0), (64, 3, 5), (4096, 2, 2 ** 64 - 1), (1 << 14, 8, 12345)])
def test_uniform_matches_oracle(word, n, L, seed):
    got = capi.sample_host("uniform", n, L, None, seed, word)
    np.testing.assert_array_equal(got, S.uniform(primes(L, word), n, seed, word))


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("n,L,sigma,seed", [(8, 1, 3.2, 1), (1024, 3, 3.2, 9), (4096, 2, 0.5, 3),
                                            (1 << 14, 8, 3.2, 77), (256, 2, 1000.0, 4)])
def test_gaussian_matches_oracle(word, n, L, sigma, seed):
    cdf = capi.gaussian_cdt(sigma)
    got = capi.sample_host("gaussian", n, L, sigma, seed, word)
    np.testing.assert_array_equal(got, S.gaussian(primes(L, word), n, cdf, seed, 1, word))


def test_degree4_samplers_match_oracle():
    """The 64-bit build's samplers at the reference's smallest ring, N = 4."""
    P = primes(3, 64)
    np.testing.assert_array_equal(capi.sample_host("uniform", 4, 3, None, 3), S.uniform(P, 4, 3, 64))
    np.testing.assert_array_equal(capi.sample_host("gaussian", 4, 3, 3.2, 3),
                                  S.gaussian(P, 4, capi.gaussian_cdt(3.2), 3, 1, 64))
    np.testing.assert_array_equal(capi.sample_host("zo", 4, 3, 127, 3), S.zo(P, 4, 127, 3, 64))
    for h in (1, 3, 4):
        np.testing.assert_array_equal(capi.sample_host("hwt", 4, 3, h, 3), S.hwt(P, 4, h, 3, 64))


def test_gaussian_noise_is_the_signed_sample():
    """ya_gaussian_noise (FastGaussianNoise::getNoise of the drop-in) returns
    the signed samples that the Gaussian sampler maps to residues, for any
    count, as a prefix of one stream."""
    cdf = capi.gaussian_cdt(3.2)
    for count in (1, 7, 4096):
        out = np.empty(count, dtype=np.int64)
        capi.call("ya_gaussian_noise", capi.ptr(out), count, 3.2, 11)
        np.testing.assert_array_equal(out, S.gaussian_signed(count, cdf, 11))
    p = [int(v) for v in capi.primes(1)]
    res = capi.sample_host("gaussian", 1024, 1, 3.2, 11)
    out = np.empty(1024, dtype=np.int64)
    capi.call("ya_gaussian_noise", capi.ptr(out), 1024, 3.2, 11)
    np.testing.assert_array_equal(np.where(out < 0, out + p[0], out).astype(np.uint64), res)


@pytest.mark.parametrize("word", [64, 32])
def test_amplifier_wraps_in_signed_type(word):
    """The amplified sample wraps in the reference's params::signed_type
    (int64_t; int32_t in the 32-bit build, params_32.hpp) before the sign test
    r < 0 ? p + r : r (details/poly.hpp:248-257).  Closed forms from C's
    conversions, 32-bit: 4 * (2^30 - 1) = 2^32 - 4 is -4 as int32_t, so p - 4;
    -4 * (2^30 - 1) is 4; 1 * 2^31 is INT32_MIN, so p + 2^31 (mod 2^32).  In
    64 bits the same products are positive / negative as integers."""
    cdf = capi.gaussian_cdt(3.2)
    r = S.gaussian_signed(4096, cdf, 5)
    p = primes(1, word)[0]
    for amp in (2 ** 30 - 1, 2 ** 31, 3 ** 20):
        out = S.gaussian([p], 4096, cdf, 5, amp, word).astype(np.uint64)
        for s_val in (4, -4, 1, -1, 0):
            idx = np.flatnonzero(r == s_val)
            if idx.size == 0:
                continue
            prod = s_val * amp
            if word == 32:
                a = ((prod + 2 ** 31) % 2 ** 32) - 2 ** 31      # the int32_t value
                want = (p + a) % 2 ** 32 if a < 0 else a
            else:
                a = ((prod + 2 ** 63) % 2 ** 64) - 2 ** 63
                want = (p + a) % 2 ** 64 if a < 0 else a
            assert set(int(v) for v in out[idx]) == {want}, (word, amp, s_val)
    if word == 32:
        idx = np.flatnonzero(r == 4)[0]
        assert int(S.gaussian([p], 4096, cdf, 5, 2 ** 30 - 1, 32)[idx]) == p - 4


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("n,L,rho,seed", [(8, 1, 127, 1), (1024, 3, 0, 2), (4096, 2, 255, 3), (1 << 14, 4, 100, 4)])
def test_zo_matches_oracle(word, n, L, rho, seed):
    got = capi.sample_host("zo", n, L, rho, seed, word)
    np.testing.assert_array_equal(got, S.zo(primes(L, word), n, rho, seed, word))


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("n,L,h,seed", [(8, 1, 1, 1), (8, 2, 8, 2), (1024, 3, 64, 3), (4096, 2, 4095, 4),
                                        (1 << 14, 2, 192, 5)])
def test_hwt_matches_oracle(word, n, L, h, seed):
    got = capi.sample_host("hwt", n, L, h, seed, word)
    np.testing.assert_array_equal(got, S.hwt(primes(L, word), n, h, seed, word))


@pytest.mark.parametrize("sigma", [0.5, 3.2, 8.0, 1000.0])
def test_gaussian_table_exact(sigma):
    """The library's long-double table within 2^-52 (relative to 2^64) of the
    exact cumulative values; monotone; saturated last entry."""
    got = [int(x) for x in capi.gaussian_cdt(sigma)]
    want = S.cdt_exact(sigma)
    assert len(got) == len(want) == int(np.ceil(13 * sigma)) + 1
    assert got[-1] == (1 << 64) - 1
    assert all(a <= b for a, b in zip(got, got[1:]))
    assert max(abs(a - b) for a, b in zip(got, want)) <= 1 << 12


def test_gaussian_table_closed_form():
    """sigma -> 0 limit and symmetry: P(X = 0) = 1 / (1 + 2 sum_{k>0} exp(-k^2/(2 s^2)))."""
    import math
    s = 0.5
    z = 1 + 2 * sum(math.exp(-k * k / (2 * s * s)) for k in range(1, 8))
    assert abs(int(capi.gaussian_cdt(s)[0]) / 2.0 ** 64 - 1 / z) < 1e-12


@pytest.mark.parametrize("word", [64, 32])
def test_distribution_properties(word):
    """Closed-form properties of each mapping at n = 2^14: uniform below p with
    mean p/2; Gaussian shared by residues with variance sigma^2; ZO nonzero with
    probability (rho + 1) / 256, +1 and -1 equally; hwt exactly hwt nonzero in
    {1, p - 1}, sign balance, shared by residues."""
    n, L = 1 << 14, 3
    ps = primes(L, word)
    u = capi.sample_host("uniform", n, L, None, 11, word)
    for j, p in enumerate(ps):
        r = u[j * n:(j + 1) * n].astype(np.float64)
        assert r.max() < p and abs(r.mean() / p - 0.5) < 5 / np.sqrt(12 * n)
    g = capi.sample_host("gaussian", n, L, 3.2, 12, word).astype(np.int64)
    signed = [np.where(g[j * n:(j + 1) * n] > p // 2, g[j * n:(j + 1) * n] - p, g[j * n:(j + 1) * n])
              for j, p in enumerate(ps)]
    assert all(np.array_equal(signed[0], x) for x in signed[1:])
    sd = signed[0].astype(np.float64)
    assert abs(sd.mean()) < 5 * 3.2 / np.sqrt(n) and abs(sd.var() / 3.2 ** 2 - 1) < 0.06
    z = capi.sample_host("zo", n, L, 127, 13, word).astype(np.int64)
    nz = np.count_nonzero(z[:n])
    assert abs(nz / n - 0.5) < 5 * 0.5 / np.sqrt(n)
    assert abs(np.count_nonzero(z[:n] == 1) / max(nz, 1) - 0.5) < 5 * 0.5 / np.sqrt(nz)
    assert set(np.unique(z[:n])) <= {0, 1, ps[0] - 1}
    for h in (1, 64, n):
        x = capi.sample_host("hwt", n, L, h, 14, word).astype(np.int64)
        for j, p in enumerate(ps):
            r = x[j * n:(j + 1) * n]
            assert np.count_nonzero(r) == h and set(np.unique(r[r != 0])) <= {1, p - 1}
            np.testing.assert_array_equal(r == 1, x[:n] == 1)


def test_hwt_positions_uniform():
    """Over many seeds every position is hit about equally often (hwt/n each)."""
    n, h, seeds = 64, 8, 600
    counts = np.zeros(n)
    for s in range(seeds):
        counts[S.hwt_positions(n, h, s)] += 1
    assert counts.sum() == h * seeds
    want = h * seeds / n
    assert np.all(np.abs(counts - want) < 6 * np.sqrt(want))
    got = capi.sample_host("hwt", n, 1, h, 3, 64)
    assert sorted(np.nonzero(got)[0].tolist()) == S.hwt_positions(n, h, 3)


def test_sampler_errors():
    lib = capi.load()
    buf = np.zeros(64, dtype=np.uint64)
    assert lib.ya_sample_hwt(capi.ptr(buf), 8, 1, 0, 1) == 1
    assert lib.ya_sample_hwt(capi.ptr(buf), 8, 1, 9, 1) == 1
    assert lib.ya_sample_zo(capi.ptr(buf), 8, 1, 256, 1) == 1
    assert lib.ya_sample_gaussian(capi.ptr(buf), 8, 1, 2e6, 1) == 1
    n = capi.ctypes.c_size_t()
    assert lib.ya_gaussian_cdt(3.2, capi.ptr(buf), 10, capi.ctypes.byref(n)) == 1   # capacity below 43
    assert lib.ya_gaussian_cdt(0.0, None, 0, capi.ctypes.byref(n)) == 1
