"""Helpers for GPU tests: device set-up and thin wrappers over the C ABI."""
import sys

import numpy as np

from yell_amd import capi

_ready = False


def device():
    """Initialise device 0 through the C ABI (raises if there is none)."""
    global _ready
    if not _ready:
        capi.call("ya_init", 0)
        _ready = True
    return capi.load()


def say(*args):
    print("[case]", *args, file=sys.stderr, flush=True)


def run_inplace(fn, host, *args):
    """Upload `host`, run ya_<fn>(dev, *args, stream=NULL), download."""
    device()
    buf = capi.DeviceBuffer.from_host(host)
    capi.call(fn, buf.ptr, *args, None)
    out = buf.download()
    buf.close()
    return out


def run_binary(fn, a, b, n, L, out_init=None):
    device()
    da = capi.DeviceBuffer.from_host(a)
    db = capi.DeviceBuffer.from_host(b)
    dc = capi.DeviceBuffer.from_host(out_init) if out_init is not None else capi.DeviceBuffer(a.size)
    capi.call(fn, dc.ptr, da.ptr, db.ptr, n, L, None)
    out = dc.download()
    for d in (da, db, dc):
        d.close()
    return out


def polymul(a, b, n, L):
    return run_binary("ya_poly_multiply", a, b, n, L)


def rand_bound(rng, n, L, mult, primes):
    out = np.empty(n * L, dtype=np.uint64)
    for r in range(L):
        out[r * n:(r + 1) * n] = rng.integers(0, int(primes[r]) * mult, size=n, dtype=np.uint64)
    return out
