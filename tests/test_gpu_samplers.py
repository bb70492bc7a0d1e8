"""Device samplers (ya_sample_*_device, ya32_sample_*_device): the same words
as the host samplers for the same seed, bit-exact, at small and full sizes
(N = 2^14, L = 8 and N = 2^16, L = 24), hwt at weight 1 and at the
degree, the Gaussian amplifier against the oracle."""
import numpy as np
import pytest

from _oracle import Oracle, sampler_oracle
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
S = sampler_oracle()
O32 = Oracle(32)

CASES = [("uniform", 8, 1, None), ("uniform", 1 << 14, 8, None), ("uniform", 1 << 16, 24, None),
         ("gaussian", 8, 1, 3.2), ("gaussian", 1 << 14, 8, 3.2), ("gaussian", 4096, 3, 1000.0),
         ("zo", 8, 1, 127), ("zo", 1 << 14, 8, 100), ("zo", 4096, 2, 255),
         ("hwt", 8, 1, 1), ("hwt", 8, 2, 8), ("hwt", 1 << 14, 8, 64), ("hwt", 1 << 14, 2, 1 << 14),
         ("hwt", 1 << 16, 3, 4096)]


def run_device(kind, n, L, param, seed, word, amplifier=1):
    device()
    buf = capi.DeviceBuffer(n * L, np.uint64 if word == 64 else np.uint32)
    capi.sample_device(kind, buf, n, L, param, seed, amplifier=amplifier, word=word)
    out = buf.download()
    buf.close()
    return out


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("kind,n,L,param", CASES)
def test_device_equals_host(word, kind, n, L, param):
    if word == 32 and n > (1 << 15):
        pytest.skip("the 32-bit build stops at 2^15")
    for seed in (1, 2 ** 63 + 5):
        say("sampler", word, kind, n, L, param, seed)
        got = run_device(kind, n, L, param, seed, word)
        np.testing.assert_array_equal(got, capi.sample_host(kind, n, L, param, seed, word))


@pytest.mark.parametrize("kind,param", [("uniform", None), ("gaussian", 3.2), ("zo", 127), ("hwt", 1),
                                        ("hwt", 4)])
def test_device_equals_host_degree4(kind, param):
    """The reference's smallest ring, N = 4 (64-bit build): the device
    samplers give the host samplers' words, which the oracle pins."""
    say("sampler N=4", kind, param)
    got = run_device(kind, 4, 3, param, 9, 64)
    np.testing.assert_array_equal(got, capi.sample_host(kind, 4, 3, param, 9))


@pytest.mark.parametrize("word", [64, 32])
@pytest.mark.parametrize("amp", [7, 2 ** 28 + 3, 2 ** 30 - 1, 2 ** 31, 2 ** 62 + 5])
def test_gaussian_amplifier(word, amp):
    """Amplified samples wrap in params::signed_type (int32_t in the 32-bit
    build), as the oracle states; large amplifiers reach the wrap."""
    n, L = 4096, 3
    say("sampler amplifier", word, amp)
    ps = [int(p) for p in capi.primes(L)] if word == 64 else [O32.prime(j) for j in range(L)]
    got = run_device("gaussian", n, L, 3.2, 21, word, amplifier=amp)
    np.testing.assert_array_equal(got, S.gaussian(ps, n, capi.gaussian_cdt(3.2), 21, amp, word))


@pytest.mark.parametrize("kind,param", [("uniform", None), ("gaussian", 3.2), ("zo", 0x7F)])
def test_many_equals_single(kind, param):
    """Polynomial k of a *_many launch = the single call with seed + k."""
    n, L, count, seed = 2048, 3, 5, 40
    say("sampler many", kind)
    device()
    buf = capi.DeviceBuffer(n * L * count)
    args = () if kind == "uniform" else ((param, 1) if kind == "gaussian" else (param,))
    capi.call("ya_sample_%s_device_many" % kind, buf.ptr, n, L, count, *args, seed, None)
    got = buf.download().reshape(count, n * L)
    buf.close()
    for k in range(count):
        np.testing.assert_array_equal(got[k], capi.sample_host(kind, n, L, param, seed + k, 64))


def test_device_sampler_errors():
    lib = device()
    buf = capi.DeviceBuffer(64)
    assert lib.ya_sample_uniform_device(buf.ptr + 8, 8, 1, 1, None) == 1      # not 16-byte aligned
    assert lib.ya_sample_hwt_device(buf.ptr, 8, 1, 0, 1, None) == 1
    assert lib.ya_sample_zo_device(buf.ptr, 8, 1, 256, 1, None) == 1
    assert lib.ya_sample_uniform_device(buf.ptr, 12, 1, 1, None) == 1
    assert lib.ya32_sample_uniform_device(buf.ptr, 1 << 16, 1, 1, None) == 1
    buf.close()


def test_many_chunks_and_extremes():
    """More polynomials than one launch's grid rows (65,535) in one call; the
    largest modulus count; the largest 64-bit degree; 32-bit at 291 moduli."""
    lib = device()
    say("sampler chunks")
    n, count = 8, 70000
    buf = capi.DeviceBuffer(n * count)
    capi.call("ya_sample_zo_device_many", buf.ptr, n, 1, count, 0x7F, 3, None)
    got = buf.download().reshape(count, n)
    for k in (0, 65534, 65535, 65536, count - 1):
        np.testing.assert_array_equal(got[k], capi.sample_host("zo", n, 1, 0x7F, 3 + k, 64))
    buf.close()
    say("sampler 1000 moduli")
    n, L = 16, 1000
    buf = capi.DeviceBuffer(n * L)
    for kind, param in (("uniform", None), ("gaussian", 3.2), ("hwt", 16)):
        capi.sample_device(kind, buf, n, L, param, 9)
        np.testing.assert_array_equal(buf.download(), capi.sample_host(kind, n, L, param, 9, 64))
    buf.close()
    say("sampler 2^17")
    n = 1 << 17
    buf = capi.DeviceBuffer(n * 2)
    for kind, param in (("uniform", None), ("hwt", 1000), ("zo", 200)):
        capi.sample_device(kind, buf, n, 2, param, 10)
        np.testing.assert_array_equal(buf.download(), capi.sample_host(kind, n, 2, param, 10, 64))
    buf.close()
    say("sampler 32-bit 291 moduli")
    buf = capi.DeviceBuffer(64 * 291, np.uint32)
    capi.sample_device("uniform", buf, 64, 291, None, 11, word=32)
    np.testing.assert_array_equal(buf.download(), capi.sample_host("uniform", 64, 291, None, 11, 32))
    buf.close()
    assert lib.ya_sample_uniform_device_many(buf.ptr, 8, 1, 0, 1, None) == 1   # count 0
