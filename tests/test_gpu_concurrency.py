"""First use of the device-resident table cache from many host threads at
once (the reference builds its tables lazily under a lock,
include/yell/details/ntt.hpp:79-104): 8 threads make their first call at a
degree and modulus range no other test uses, concurrently; every result
equals the oracle's and the resident tables equal the oracle's tables."""
import threading

import numpy as np
import pytest

from _oracle import Oracle
from gpu_util import device, rand_bound, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
orc = Oracle()


@pytest.mark.parametrize("n,L", [(1 << 5, 300), (1 << 11, 60)])
def test_concurrent_first_use_of_tables(n, L):
    device()
    P = [orc.prime(i) for i in range(L)]
    nthreads = 8
    inputs = [rand_bound(np.random.default_rng(50 + k), n, L, 4, P) for k in range(nthreads)]
    results, errors = [None] * nthreads, []
    start = threading.Barrier(nthreads)

    def work(k):
        try:
            buf = capi.DeviceBuffer.from_host(inputs[k])
            start.wait()
            capi.call("ya_poly_forward", buf.ptr, n, L, None)   # builds the (n, P[0..L)) tables on first use
            capi.call("ya_stream_synchronize", None)
            results[k] = buf.download()
            buf.close()
        except Exception as ex:  # reported below, in the main thread
            errors.append(ex)

    say("concurrent first use", n, L, nthreads, "threads")
    threads = [threading.Thread(target=work, args=(k,)) for k in range(nthreads)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k in range(nthreads):
        np.testing.assert_array_equal(results[k], orc.poly_forward(inputs[k], n, L), err_msg="thread %d" % k)
    for cm in (0, L // 2, L - 1):
        for got, want in zip(capi.table_download(n, cm), orc.tables(n, cm)):
            np.testing.assert_array_equal(got, want)
