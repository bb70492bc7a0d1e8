"""The looping fused kernel (k_polymul_loop<14,4>): ya_poly_multiply_many at
N = 2^14 with at least two residues per workgroup of a one-per-CU grid runs
it; every polynomial is checked against the oracle's poly-multiplication
(the reference's a.forward(); b.forward(); c = a * b; c.backward()), and
the whole output against the one-workgroup-per-residue kernel (dual forward
off), at residue counts that fill the last round exactly and that leave it
ragged, with 1, 3 and 8 moduli."""
import numpy as np
import pytest

from _oracle import Oracle
from gpu_util import device, say
from yell_amd import capi

pytestmark = pytest.mark.gpu
orc = Oracle()
N = 1 << 14


@pytest.mark.parametrize("L,count", [(8, 64), (8, 67), (3, 180), (1, 523)])
def test_loop_kernel_matches_oracle_and_grid_kernel(L, count):
    say("polymul loop", N, L, count)
    lib = device()
    w = N * L
    a = np.concatenate([capi.sample_uniform(N, L, 8100 + k) for k in range(count)])
    b = np.concatenate([capi.sample_uniform(N, L, 8500 + k) for k in range(count)])
    b[:2] = b[2:4]  # repeated words
    da, db = capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b)
    dc = capi.DeviceBuffer(w * count)
    capi.call("ya_ntt_init_tables", N, L)
    saved = (lib.ya_tune_get(1, 14), lib.ya_tune_get(3, 14))
    try:
        capi.call("ya_tune_set", 1, 14, 4)
        capi.call("ya_tune_set", 3, 14, 1)   # dual forward: the looping kernel at this count
        capi.call("ya_poly_multiply_many", dc.ptr, da.ptr, db.ptr, N, L, count, None)
        capi.call("ya_device_synchronize")
        looped = dc.download()
        capi.call("ya_tune_set", 3, 14, 0)   # one workgroup per residue
        capi.call("ya_poly_multiply_many", dc.ptr, da.ptr, db.ptr, N, L, count, None)
        capi.call("ya_device_synchronize")
        grid = dc.download()
    finally:
        capi.call("ya_tune_set", 1, 14, saved[0])
        capi.call("ya_tune_set", 3, 14, saved[1])
    np.testing.assert_array_equal(looped, grid)
    for k in range(count):
        sl = slice(k * w, (k + 1) * w)
        np.testing.assert_array_equal(looped[sl], orc.polymul(a[sl], b[sl], N, L), err_msg="poly %d" % k)
    for d in (da, db, dc):
        d.close()
