"""ctypes binding of include/yell_amd.h (the engine's C ABI).

Loads yell_amd/lib/libyell_amd.so built in-tree; raises ImportError-style
RuntimeError when it is missing (there is no fallback path).  Every binding
keeps the C signature: device buffers are integers (device addresses), host
buffers are numpy uint64 arrays.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libyell_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "yell_amd.h")

u64 = ctypes.c_uint64
sz = ctypes.c_size_t
vp = ctypes.c_void_p
i32 = ctypes.c_int
u64p = ctypes.POINTER(ctypes.c_uint64)

# name -> (restype, argtypes)
_SIGS = {
    "ya_status_string": (ctypes.c_char_p, [i32]),
    "ya_last_error": (ctypes.c_char_p, []),
    "ya_init": (i32, [i32]),
    "ya_device_count": (i32, [ctypes.POINTER(i32)]),
    "ya_device_synchronize": (i32, []),
    "ya_num_moduli": (sz, []),
    "ya_prime": (u64, [sz]),
    "ya_barrett_const": (u64, [sz]),
    "ya_root": (u64, [sz]),
    "ya_ntt_init_tables": (i32, [sz, sz]),
    "ya_ntt_table_host": (i32, [sz, sz, vp, vp, vp, vp, vp]),
    "ya_ntt_table_download": (i32, [sz, sz, vp, vp, vp, vp, vp]),
    "ya_malloc": (i32, [ctypes.POINTER(vp), sz]),
    "ya_free": (i32, [vp]),
    "ya_memset_zero": (i32, [vp, sz, vp]),
    "ya_memcpy_htod": (i32, [vp, vp, sz, vp]),
    "ya_memcpy_dtoh": (i32, [vp, vp, sz, vp]),
    "ya_memcpy_dtod": (i32, [vp, vp, sz, vp]),
    "ya_host_alloc_pinned": (i32, [ctypes.POINTER(vp), sz]),
    "ya_host_free_pinned": (i32, [vp]),
    "ya_stream_create": (i32, [ctypes.POINTER(vp)]),
    "ya_stream_destroy": (i32, [vp]),
    "ya_stream_synchronize": (i32, [vp]),
    "ya_event_create": (i32, [ctypes.POINTER(vp)]),
    "ya_event_destroy": (i32, [vp]),
    "ya_event_record": (i32, [vp, vp]),
    "ya_event_synchronize": (i32, [vp]),
    "ya_event_elapsed_ms": (i32, [ctypes.POINTER(ctypes.c_float), vp, vp]),
    "ya_ntt_forward": (i32, [vp, sz, sz, vp]),
    "ya_ntt_forward_lazy": (i32, [vp, sz, sz, vp]),
    "ya_ntt_backward": (i32, [vp, sz, sz, vp]),
    "ya_negacylic_forward_lazy": (i32, [vp, sz, vp, vp, u64]),
    "ya_negacylic_backward_lazy": (i32, [vp, sz, vp, vp, u64]),
    "ya_poly_forward": (i32, [vp, sz, sz, vp]),
    "ya_poly_forward_lazy": (i32, [vp, sz, sz, vp]),
    "ya_poly_backward": (i32, [vp, sz, sz, vp]),
    "ya_poly_mul": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya_poly_add": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya_poly_sub": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya_poly_negate": (i32, [vp, sz, sz, vp]),
    "ya_poly_add_product_of": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya_poly_multiply": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya_poly_forward_many": (i32, [vp, sz, sz, sz, vp]),
    "ya_pair_forward": (i32, [i32]),
    "ya_batch_strided": (i32, [i32]),
    "ya_batch_plan_is_strided": (i32, [vp, ctypes.POINTER(i32)]),
    "ya_poly_forward_lazy_many": (i32, [vp, sz, sz, sz, vp]),
    "ya_poly_backward_many": (i32, [vp, sz, sz, sz, vp]),
    "ya_poly_add_products": (i32, [vp, vp, vp, sz, sz, sz, vp]),
    "ya_poly_add_products_backward": (i32, [vp, vp, vp, sz, sz, sz, vp]),
    "ya_poly_sum_products_backward": (i32, [vp, vp, vp, sz, sz, sz, vp]),
    "ya_poly_multiply_many": (i32, [vp, vp, vp, sz, sz, sz, vp]),
    "ya_poly_multiply_noalloc": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya_sample_uniform": (i32, [vp, sz, sz, u64]),
    "ya_sample_gaussian": (i32, [vp, sz, sz, ctypes.c_double, u64]),
    "ya_sample_zo": (i32, [vp, sz, sz, ctypes.c_uint32, u64]),
    "ya_sample_hwt": (i32, [vp, sz, sz, ctypes.c_uint32, u64]),
    "ya_graph_begin": (i32, [vp]),
    "ya_graph_end": (i32, [vp, ctypes.POINTER(vp)]),
    "ya_graph_launch": (i32, [vp, vp]),
    "ya_graph_destroy": (i32, [vp]),
    "ya_gaussian_noise": (i32, [vp, sz, ctypes.c_double, u64]),
    "ya_gaussian_cdt": (i32, [ctypes.c_double, vp, sz, ctypes.POINTER(ctypes.c_size_t)]),
    "ya_sample_uniform_device": (i32, [vp, sz, sz, u64, vp]),
    "ya_sample_gaussian_device": (i32, [vp, sz, sz, ctypes.c_double, u64, u64, vp]),
    "ya_sample_zo_device": (i32, [vp, sz, sz, ctypes.c_uint32, u64, vp]),
    "ya_sample_hwt_device": (i32, [vp, sz, sz, ctypes.c_uint32, u64, vp]),
    "ya_sample_uniform_device_many": (i32, [vp, sz, sz, sz, u64, vp]),
    "ya_sample_gaussian_device_many": (i32, [vp, sz, sz, sz, ctypes.c_double, u64, u64, vp]),
    "ya_sample_zo_device_many": (i32, [vp, sz, sz, sz, ctypes.c_uint32, u64, vp]),
    "ya_poly_mul_many": (i32, [vp, vp, vp, sz, sz, sz, i32, vp]),
    "ya_poly_add_many": (i32, [vp, vp, vp, sz, sz, sz, i32, vp]),
    "ya_poly_sub_many": (i32, [vp, vp, vp, sz, sz, sz, i32, vp]),
    "ya_poly_add_product_of_many": (i32, [vp, vp, vp, sz, sz, sz, i32, vp]),
    "ya_poly_negate_many": (i32, [vp, sz, sz, sz, vp]),
    "ya_poly_forward_lazy_add_product_many": (i32, [vp, vp, vp, vp, sz, sz, sz, vp]),
    "ya_poly_add_product_backward_many": (i32, [vp, vp, vp, vp, sz, sz, sz, vp]),
    "ya_pointer_is_device": (i32, [vp, ctypes.POINTER(i32)]),
    "ya_probe_valu": (i32, [i32, ctypes.POINTER(ctypes.c_double)]),
    "ya_crt_limbs": (i32, [sz, ctypes.POINTER(ctypes.c_size_t)]),
    "ya_split_policy": (i32, [i32]),
    "ya32_num_moduli": (sz, []),
    "ya32_prime": (ctypes.c_uint32, [sz]),
    "ya32_barrett_const": (ctypes.c_uint32, [sz]),
    "ya32_root": (ctypes.c_uint32, [sz]),
    "ya32_ntt_table_host": (i32, [sz, sz, vp, vp, vp, vp, vp]),
    "ya32_ntt_init_tables": (i32, [sz, sz]),
    "ya32_ntt_forward": (i32, [vp, sz, sz, vp]),
    "ya32_ntt_forward_lazy": (i32, [vp, sz, sz, vp]),
    "ya32_ntt_backward": (i32, [vp, sz, sz, vp]),
    "ya32_poly_forward": (i32, [vp, sz, sz, vp]),
    "ya32_poly_forward_lazy": (i32, [vp, sz, sz, vp]),
    "ya32_poly_backward": (i32, [vp, sz, sz, vp]),
    "ya32_poly_mul": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya32_poly_add": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya32_poly_sub": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya32_poly_negate": (i32, [vp, sz, sz, vp]),
    "ya32_poly_add_product_of": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya32_poly_multiply": (i32, [vp, vp, vp, sz, sz, vp]),
    "ya32_poly_multiply_many": (i32, [vp, vp, vp, sz, sz, sz, vp]),
    "ya32_sample_uniform": (i32, [vp, sz, sz, u64]),
    "ya32_sample_gaussian": (i32, [vp, sz, sz, ctypes.c_double, u64]),
    "ya32_sample_zo": (i32, [vp, sz, sz, ctypes.c_uint32, u64]),
    "ya32_sample_hwt": (i32, [vp, sz, sz, ctypes.c_uint32, u64]),
    "ya32_sample_uniform_device": (i32, [vp, sz, sz, u64, vp]),
    "ya32_sample_gaussian_device": (i32, [vp, sz, sz, ctypes.c_double, u64, u64, vp]),
    "ya32_sample_zo_device": (i32, [vp, sz, sz, ctypes.c_uint32, u64, vp]),
    "ya32_sample_hwt_device": (i32, [vp, sz, sz, ctypes.c_uint32, u64, vp]),
    "ya_poly_crt_lift": (i32, [vp, vp, sz, sz, vp]),
    "ya_crt_constants": (i32, [sz, sz, vp, vp, vp, ctypes.POINTER(ctypes.c_size_t)]),
    "ya_batch_plan_create": (i32, [ctypes.POINTER(vp), vp, sz]),
    "ya_batch_plan_run": (i32, [vp, vp]),
    "ya_batch_plan_run_graph": (i32, [vp, vp]),
    "ya_batch_plan_info": (i32, [vp, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz),
                                 ctypes.POINTER(i32)]),
    "ya_batch_force_form": (i32, [i32]),
    "ya_autotune_mac": (i32, [sz, sz, sz, ctypes.POINTER(ctypes.c_double)]),
    "ya_batch_plan_destroy": (i32, [vp]),
    "ya_batch_polymul": (i32, [vp, sz, vp]),
    "ya_autotune": (i32, [i32, i32, ctypes.POINTER(ctypes.c_double)]),
    "ya_tune_get": (i32, [i32, i32]),
    "ya_tune_set": (i32, [i32, i32, i32]),
    "ya_server_create": (i32, [ctypes.POINTER(vp), vp]),
    "ya_server_submit": (i32, [vp, vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(u64)]),
    "ya_server_wait": (i32, [vp, u64, i32, ctypes.POINTER(ctypes.c_double)]),
    "ya_server_poll": (i32, [vp, u64, ctypes.POINTER(i32)]),
    "ya_server_get_stats": (i32, [vp, vp]),
    "ya_server_reset_stats": (i32, [vp]),
    "ya_server_destroy": (i32, [vp]),
    "ya_server_run_poisson": (i32, [vp, vp, vp]),
}


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_batch_jobs", ctypes.c_uint32),
                ("max_inflight", ctypes.c_uint32), ("max_pending", ctypes.c_uint32),
                ("max_moduli", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "rejected", "batches",
                                              "max_pending", "arena_bytes_in_use")] + \
               [(k, ctypes.c_double) for k in ("mean_batch_jobs", "p50_us", "p90_us", "p99_us",
                                              "max_us", "mean_us", "p50_queue_us", "p99_queue_us",
                                              "p50_service_us", "p99_service_us",
                                              "p50_batch_gpu_us", "p99_batch_gpu_us", "p99_batch_d2h_us",
                                              "p99_batch_launch_us", "p50_batch_host_us", "p99_batch_host_us")]


class PoissonSpec(ctypes.Structure):
    _fields_ = [("rate_per_s", ctypes.c_double), ("duration_s", ctypes.c_double),
                ("nshapes", ctypes.c_uint32), ("degrees", ctypes.POINTER(ctypes.c_uint32)),
                ("nmoduli", ctypes.POINTER(ctypes.c_uint32)), ("pool", ctypes.c_uint32),
                ("verify", ctypes.c_uint32), ("seed", ctypes.c_uint64)]


class ServeReport(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("offered", "completed", "rejected", "failed",
                                              "mismatches", "batches", "max_pending",
                                              "client_stalls")] + \
               [(k, ctypes.c_double) for k in ("wall_s", "throughput_per_s", "p50_us", "p90_us",
                                              "p99_us", "max_us", "mean_us", "mean_batch_jobs",
                                              "p50_queue_us", "p99_queue_us", "p50_service_us",
                                              "p99_service_us",
                                              "p50_batch_gpu_us", "p99_batch_gpu_us", "p99_batch_d2h_us",
                                              "p99_batch_launch_us", "p50_batch_host_us", "p99_batch_host_us")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PolymulJob(ctypes.Structure):
    """struct ya_polymul_job."""
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("degree", ctypes.c_uint32), ("nmoduli", ctypes.c_uint32)]


class YaError(RuntimeError):
    def __init__(self, code, fn, detail):
        super().__init__("%s failed: %s (%d): %s" % (fn, STATUS.get(code, "?"), code, detail))
        self.code = code


STATUS = {0: "YA_OK", 1: "YA_ERR_INVALID", 2: "YA_ERR_NO_DEVICE", 3: "YA_ERR_ALLOC",
          4: "YA_ERR_LAUNCH", 5: "YA_ERR_UNSUPPORTED", 6: "YA_ERR_REJECTED", 7: "YA_ERR_TIMEOUT",
          8: "YA_ERR_MISMATCH"}

_lib = None


def declared_symbols(header=HEADER):
    """Function names declared in include/*.h (for the export test)."""
    names = set()
    for h in ([header] if isinstance(header, str) else header):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(ya_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def load(path=LIB_PATH):
    """Load the HIP engine.  Raises RuntimeError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libyell_amd.so not built at %s (run python3 yell_amd/build.py)" % path)
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, fn="call"):
    if rc != 0:
        detail = _lib.ya_last_error().decode() if _lib is not None else ""
        raise YaError(rc, fn, detail)
    return rc


def call(name, *args):
    lib = load()
    return check(getattr(lib, name)(*args), name)


def ptr(a):
    """Host numpy array -> void* for ctypes."""
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceBuffer:
    """A device allocation of uint64 words (uint32 with dtype=np.uint32, the
    32-bit moduli build), freed on close / garbage collection."""

    def __init__(self, nwords, dtype=np.uint64):
        self.nwords = int(nwords)
        self.dtype = np.dtype(dtype)
        p = ctypes.c_void_p()
        call("ya_malloc", ctypes.byref(p), self.nwords * self.dtype.itemsize)
        self.ptr = p.value or 0

    @classmethod
    def from_host(cls, arr, stream=None):
        arr = np.asarray(arr)
        dt = np.uint32 if arr.dtype == np.uint32 else np.uint64
        arr = np.ascontiguousarray(arr, dtype=dt)
        b = cls(arr.size, dt)
        b.upload(arr, stream)
        return b

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.size <= self.nwords
        call("ya_memcpy_htod", self.ptr, ptr(arr), arr.nbytes, stream)

    def download(self, nwords=None, stream=None):
        n = self.nwords if nwords is None else nwords
        out = np.empty(n, dtype=self.dtype)
        call("ya_memcpy_dtoh", ptr(out), self.ptr, out.nbytes, stream)
        return out

    def close(self):
        if self.ptr and _lib is not None:
            _lib.ya_free(self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def primes(n):
    lib = load()
    return np.array([lib.ya_prime(i) for i in range(n)], dtype=np.uint64)


def sample_uniform(degree, nmoduli, seed):
    out = np.empty(degree * nmoduli, dtype=np.uint64)
    call("ya_sample_uniform", ptr(out), degree, nmoduli, seed)
    return out


def sample_gaussian(degree, nmoduli, sigma, seed):
    out = np.empty(degree * nmoduli, dtype=np.uint64)
    call("ya_sample_gaussian", ptr(out), degree, nmoduli, sigma, seed)
    return out


def gaussian_cdt(sigma):
    """The Gaussian samplers' cumulative table (uint64, ceil(13 sigma) + 1 entries)."""
    n = ctypes.c_size_t()
    call("ya_gaussian_cdt", sigma, None, 0, ctypes.byref(n))
    out = np.empty(n.value, dtype=np.uint64)
    call("ya_gaussian_cdt", sigma, ptr(out), out.size, ctypes.byref(n))
    return out


def sample_host(kind, degree, nmoduli, param, seed, word=64):
    """Host sampler `kind` ('uniform', 'gaussian', 'zo', 'hwt') of the 64- or
    32-bit build; `param` is sigma, rho or hwt (ignored for uniform)."""
    dt = np.uint64 if word == 64 else np.uint32
    pre = "ya_" if word == 64 else "ya32_"
    out = np.empty(degree * nmoduli, dtype=dt)
    args = () if kind == "uniform" else (param,)
    call(pre + "sample_" + kind, ptr(out), degree, nmoduli, *args, seed)
    return out


def sample_device(kind, buf, degree, nmoduli, param, seed, amplifier=1, stream=None, word=64):
    """Device sampler `kind` into DeviceBuffer `buf` (the same words as
    sample_host with the same seed; gaussian takes `amplifier`)."""
    pre = "ya_" if word == 64 else "ya32_"
    if kind == "uniform":
        args = ()
    elif kind == "gaussian":
        args = (param, amplifier)
    else:
        args = (param,)
    call(pre + "sample_" + kind + "_device", buf.ptr, degree, nmoduli, *args, seed, stream)


def table_host(degree, cm):
    arrs = [np.empty(degree, dtype=np.uint64) for _ in range(4)]
    sc = np.empty(4, dtype=np.uint64)
    call("ya_ntt_table_host", degree, cm, *[ptr(a) for a in arrs], ptr(sc))
    return arrs + [sc]


def table_download(degree, cm):
    arrs = [np.empty(degree, dtype=np.uint64) for _ in range(4)]
    sc = np.empty(4, dtype=np.uint64)
    call("ya_ntt_table_download", degree, cm, *[ptr(a) for a in arrs], ptr(sc))
    return arrs + [sc]
