"""The looping fused kernel (k_polymul_loop<14,4>): ya_poly_multiply_many at
N = 2^14 with at least two residues per workgroup of a one-per-CU grid runs
it (and a grid-form plan over separately allocated polynomials runs its
residue-table form, k_polymul_tab_loop); every polynomial is checked
against the oracle's poly-multiplication (the reference's a.forward();
b.forward(); c = a * b; c.backward()), and the whole output against the
one-workgroup-per-residue kernel (dual forward off), at residue counts that
fill the last round exactly and that leave it ragged, with 1, 3 and 8
moduli."""
import ctypes

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


@pytest.mark.parametrize("count", [64, 67])
def test_table_loop_kernel_matches_oracle_and_grid_kernel(count):
    """A grid-form plan over separately allocated polynomials reads the
    residue table: with the dual forward on it runs k_polymul_tab_loop
    (at least two residues per CU), with it off k_polymul_tab; both give
    the oracle's words for every polynomial."""
    L = 8
    say("polymul table loop", N, L, count)
    lib = device()
    bufs = []
    for k in range(count):
        a = capi.sample_uniform(N, L, 9100 + k)
        b = capi.sample_uniform(N, L, 9500 + k)
        bufs.append((a, b, capi.DeviceBuffer.from_host(a), capi.DeviceBuffer.from_host(b), capi.DeviceBuffer(N * L)))
    capi.call("ya_ntt_init_tables", N, L)
    arr = (capi.PolymulJob * count)()
    for k, (_, _, da, db, dc) in enumerate(bufs):
        arr[k] = capi.PolymulJob(da.ptr, db.ptr, dc.ptr, N, L)
    saved = (lib.ya_tune_get(1, 14), lib.ya_tune_get(3, 14))
    outs = {}
    capi.call("ya_batch_force_form", 2)
    try:
        capi.call("ya_tune_set", 1, 14, 4)
        for dual in (1, 0):
            capi.call("ya_tune_set", 3, 14, dual)
            plan = ctypes.c_void_p()
            capi.call("ya_batch_plan_create", ctypes.byref(plan), ctypes.cast(arr, ctypes.c_void_p), count)
            form = ctypes.c_int()
            capi.call("ya_batch_plan_info", plan, None, None, None, ctypes.byref(form))
            assert form.value == 2
            capi.call("ya_batch_plan_run", plan, None)
            capi.call("ya_device_synchronize")
            outs[dual] = [dc.download() for _, _, _, _, dc in bufs]
            lib.ya_batch_plan_destroy(plan)
    finally:
        capi.call("ya_batch_force_form", 0)
        capi.call("ya_tune_set", 1, 14, saved[0])
        capi.call("ya_tune_set", 3, 14, saved[1])
    for k, (a, b, da, db, dc) in enumerate(bufs):
        np.testing.assert_array_equal(outs[1][k], outs[0][k], err_msg="poly %d" % k)
        np.testing.assert_array_equal(outs[1][k], orc.polymul(a, b, N, L), err_msg="poly %d" % k)
        for d in (da, db, dc):
            d.close()
