"""ctypes access to the CPU checker (oracle/): the C restatement, the GMP
schoolbook and, where it was built, the reference compiled from its own
sources (oracle/_ref).  Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
ORACLE32 = os.path.join(ROOT, "oracle", "lib", "liboracle32.so")
ORACLE_GMP = os.path.join(ROOT, "oracle", "lib", "liboracle_gmp.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libyellref.so")
REF32 = os.path.join(ROOT, "oracle", "_ref", "libyellref32.so")
REF_GMP = os.path.join(ROOT, "oracle", "_ref", "libyellref_gmp.so")

u64, sz, vp, i32 = ctypes.c_uint64, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


class Oracle:
    """The C restatement (oracle/lib/liboracle.so); word=32 selects the
    32-bit moduli build (liboracle32.so, or32_* names, uint32 words)."""

    def __init__(self, word=64):
        path = ORACLE if word == 64 else ORACLE32
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        self.word = word
        self.dt = np.uint64 if word == 64 else np.uint32
        w = u64 if word == 64 else ctypes.c_uint32
        pre = "or_" if word == 64 else "or32_"
        sigs = {
            "prime": (w, [sz]), "barrett_const": (w, [sz]), "root": (w, [sz]),
            "barrett_hl": (w, [w, w, sz]), "shoupify": (w, [w, sz]),
            "mulmod": (w, [w, w, sz]), "mulmod_shoup": (w, [w, w, w, sz]),
            "addmod": (w, [w, w, sz]), "submod": (w, [w, w, sz]),
            "muladd": (w, [w, w, w, sz]),
            "negacyclic_forward_lazy": (i32, [vp, sz, vp, vp, w]),
            "negacyclic_backward_lazy": (i32, [vp, sz, vp, vp, w]),
            "tables_init": (i32, [sz, sz, vp, vp, vp, vp, vp]),
            "forward": (i32, [vp, sz, sz]), "forward_lazy": (i32, [vp, sz, sz]),
            "backward": (i32, [vp, sz, sz]),
            "poly_forward": (i32, [vp, sz, sz]), "poly_forward_lazy": (i32, [vp, sz, sz]),
            "poly_backward": (i32, [vp, sz, sz]),
            "poly_mul": (None, [vp, vp, vp, sz, sz]), "poly_add": (None, [vp, vp, vp, sz, sz]),
            "poly_sub": (None, [vp, vp, vp, sz, sz]), "poly_negate": (None, [vp, sz, sz]),
            "poly_add_product_of": (None, [vp, vp, vp, sz, sz]),
            "polymul": (i32, [vp, vp, vp, sz, sz]),
            "schoolbook": (None, [vp, vp, vp, sz, sz]),
        }
        lib = _bind(ctypes.CDLL(path), {pre + k: v for k, v in sigs.items()})
        self.raw = lib
        # unprefixed access: self.lib.or_mulmod(...) style names map to this build
        self.lib = _Prefixed(lib, pre)

    def _arr(self, x):
        return np.ascontiguousarray(x, dtype=self.dt)

    def prime(self, cm):
        return int(self.lib.or_prime(cm))

    def tables(self, n, cm):
        arrs = [np.empty(n, dtype=self.dt) for _ in range(4)]
        sc = np.empty(4, dtype=self.dt)
        assert self.lib.or_tables_init(n, cm, *[_p(a) for a in arrs], _p(sc)) == 0
        return arrs + [sc]

    def _inplace(self, fn, x, *args):
        y = self._arr(x).copy()
        assert getattr(self.lib, fn)(_p(y), *args) == 0
        return y

    def forward(self, x, n, cm):
        return self._inplace("or_forward", x, n, cm)

    def forward_lazy(self, x, n, cm):
        return self._inplace("or_forward_lazy", x, n, cm)

    def backward(self, x, n, cm):
        return self._inplace("or_backward", x, n, cm)

    def poly_forward(self, x, n, L):
        return self._inplace("or_poly_forward", x, n, L)

    def poly_forward_lazy(self, x, n, L):
        return self._inplace("or_poly_forward_lazy", x, n, L)

    def poly_backward(self, x, n, L):
        return self._inplace("or_poly_backward", x, n, L)

    def negacyclic_forward_lazy(self, x, n, w, ws, p):
        y = self._arr(x).copy()
        assert self.lib.or_negacyclic_forward_lazy(_p(y), n, _p(w), _p(ws), p) == 0
        return y

    def negacyclic_backward_lazy(self, x, n, w, ws, p):
        y = self._arr(x).copy()
        assert self.lib.or_negacyclic_backward_lazy(_p(y), n, _p(w), _p(ws), p) == 0
        return y

    def _binary(self, fn, a, b, n, L):
        out = np.empty(n * L, dtype=self.dt)
        a, b = self._arr(a), self._arr(b)
        getattr(self.lib, fn)(_p(out), _p(a), _p(b), n, L)
        return out

    def poly_mul(self, a, b, n, L):
        return self._binary("or_poly_mul", a, b, n, L)

    def poly_add(self, a, b, n, L):
        return self._binary("or_poly_add", a, b, n, L)

    def poly_sub(self, a, b, n, L):
        return self._binary("or_poly_sub", a, b, n, L)

    def poly_negate(self, a, n, L):
        y = self._arr(a).copy()
        self.lib.or_poly_negate(_p(y), n, L)
        return y

    def poly_add_product_of(self, acc, a, b, n, L):
        y = self._arr(acc).copy()
        a, b = self._arr(a), self._arr(b)
        self.lib.or_poly_add_product_of(_p(y), _p(a), _p(b), n, L)
        return y

    def polymul(self, a, b, n, L):
        out = np.empty(n * L, dtype=self.dt)
        a, b = self._arr(a), self._arr(b)
        assert self.lib.or_polymul(_p(out), _p(a), _p(b), n, L) == 0
        return out

    def schoolbook(self, a, b, n, cm):
        out = np.empty(n, dtype=self.dt)
        a, b = self._arr(a), self._arr(b)
        self.lib.or_schoolbook(_p(out), _p(a), _p(b), n, cm)
        return out


class _Prefixed:
    """Maps or_<name> attribute access to <prefix><name> of a library."""

    def __init__(self, lib, prefix):
        self._lib, self._pre = lib, prefix

    def __getattr__(self, name):
        if name.startswith("or_"):
            name = self._pre + name[3:]
        return getattr(self._lib, name)


class GmpOracle:
    def __init__(self):
        if not os.path.exists(ORACLE_GMP):
            raise RuntimeError("GMP oracle not built (gmp.h absent)")
        self.lib = _bind(ctypes.CDLL(ORACLE_GMP), {
            "or_gmp_schoolbook": (i32, [vp, vp, vp, sz, u64])})

    def schoolbook(self, a, b, n, p):
        out = np.empty(n, dtype=np.uint64)
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b, dtype=np.uint64)
        assert self.lib.or_gmp_schoolbook(_p(out), _p(a), _p(b), n, p) == 0
        return out


class RefLib:
    """The reference's own lib/{ntt,ops,math,params}.cpp compiled in oracle/_ref;
    word=32: its 32-bit moduli build (libyellref32.so)."""

    def __init__(self, word=64):
        path = REF if word == 64 else REF32
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref not built (reference checkout absent)")
        self.dt = np.uint64 if word == 64 else np.uint32
        w = u64 if word == 64 else ctypes.c_uint32
        self.lib = _bind(ctypes.CDLL(path), {
            "yref_forward": (i32, [i32, vp, sz]), "yref_forward_lazy": (i32, [i32, vp, sz]),
            "yref_backward": (i32, [i32, vp, sz]),
            "yref_tables": (i32, [i32, sz, vp, vp, vp, vp, vp]),
            "yref_init_tables": (i32, [i32, sz]),
            "yref_polymul": (i32, [i32, vp, vp, vp, sz]),
            "yref_bench_polymul": (ctypes.c_double, [i32, vp, vp, sz, i32, i32]),
            "yref_mulmod": (w, [w, w, sz]), "yref_mulmod_shoup": (w, [w, w, w, sz]),
            "yref_addmod": (w, [w, w, sz]), "yref_submod": (w, [w, w, sz]),
            "yref_muladd": (w, [w, w, w, sz]), "yref_shoupify": (w, [w, sz]),
            "yref_barrett": (w, [w, w, sz]),
            "yref_rlwe": (i32, [i32, sz] + [vp] * 8),
            "yref_bench_rlwe": (ctypes.c_double, [i32, sz] + [vp] * 7 + [i32, i32]),
        })

    def _inplace(self, fn, x, n, cm):
        y = np.ascontiguousarray(x, dtype=self.dt).copy()
        assert getattr(self.lib, fn)(n.bit_length() - 1, _p(y), cm) == 0
        return y

    def forward(self, x, n, cm):
        return self._inplace("yref_forward", x, n, cm)

    def forward_lazy(self, x, n, cm):
        return self._inplace("yref_forward_lazy", x, n, cm)

    def backward(self, x, n, cm):
        return self._inplace("yref_backward", x, n, cm)

    def tables(self, n, cm):
        arrs = [np.empty(n, dtype=self.dt) for _ in range(4)]
        sc = np.empty(4, dtype=self.dt)
        assert self.lib.yref_tables(n.bit_length() - 1, cm, *[_p(a) for a in arrs], _p(sc)) == 0
        return arrs + [sc]

    def polymul(self, a, b, n, L):
        out = np.empty(n * L, dtype=self.dt)
        a = np.ascontiguousarray(a, dtype=self.dt)
        b = np.ascontiguousarray(b, dtype=self.dt)
        assert self.lib.yref_polymul(n.bit_length() - 1, _p(out), _p(a), _p(b), L) == 0
        return out


    def rlwe(self, n, L, sx, e0, pk_ax, msg, u, e1, e2):
        """example.cpp:12-58 through the reference's sources:
        (sx in NTT form, pk_bx, ct_bx, ct_ax, plain)."""
        c = lambda x: np.ascontiguousarray(x, dtype=self.dt).copy()
        sx, bx, ax, m, u, cb, ca = c(sx), c(e0), c(pk_ax), c(msg), c(u), c(e1), c(e2)
        plain = np.empty(n * L, dtype=self.dt)
        assert self.lib.yref_rlwe(n.bit_length() - 1, L, _p(sx), _p(bx), _p(ax), _p(m), _p(u), _p(cb), _p(ca),
                                  _p(plain)) == 0
        return sx, bx, cb, ca, plain

    def bench_rlwe(self, n, L, sxf, pk_bx, pk_ax, msg, u, e1, e2, threads, reps):
        """Seconds for `reps` encrypt + decrypt rounds on each of `threads` threads."""
        c = lambda x: np.ascontiguousarray(x, dtype=self.dt)
        args = [c(x) for x in (sxf, pk_bx, pk_ax, msg, u, e1, e2)]
        return self.lib.yref_bench_rlwe(n.bit_length() - 1, L, *[_p(a) for a in args], threads, reps)


class RefGmpLib:
    """The reference's CRT lift (lib/gmp.cpp, gmp.hpp poly2mpz) compiled in oracle/_ref."""

    def __init__(self):
        if not os.path.exists(REF_GMP):
            raise RuntimeError("oracle/_ref/libyellref_gmp.so not built")
        self.lib = _bind(ctypes.CDLL(REF_GMP), {
            "yrefg_poly2mpz": (i32, [vp, i32, sz, vp, sz]),
            "yrefg_constants": (i32, [sz, sz, vp, vp, vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]),
            "yrefg_bench_poly2mpz": (ctypes.c_double, [vp, i32, sz, i32]),
        })

    def poly2mpz(self, res, n, L, limbs):
        out = np.empty((n, limbs), dtype=np.uint64)
        res = np.ascontiguousarray(res, dtype=np.uint64)
        rc = self.lib.yrefg_poly2mpz(_p(res), n.bit_length() - 1, L, _p(out), limbs)
        assert rc == 0, rc
        return out

    def constants(self, L, limbs):
        q = np.empty(limbs, dtype=np.uint64)
        m = np.empty(2, dtype=np.uint64)
        lift = np.empty((L, limbs), dtype=np.uint64)
        shift, bits = sz(), sz()
        assert self.lib.yrefg_constants(L, limbs, _p(q), _p(m), _p(lift), ctypes.byref(shift),
                                        ctypes.byref(bits)) == 0
        return q, m, lift, shift.value, bits.value

    def bench(self, res, n, L, reps):
        res = np.ascontiguousarray(res, dtype=np.uint64)
        return self.lib.yrefg_bench_poly2mpz(_p(res), n.bit_length() - 1, L, reps)


def crt_oracle():
    """oracle/crt_oracle.py (Python-integer restatement of the CRT lift)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("crt_oracle", os.path.join(ROOT, "oracle", "crt_oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def sampler_oracle():
    """oracle/sampler_oracle.py (the samplers over numpy's Philox)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("sampler_oracle",
                                                  os.path.join(ROOT, "oracle", "sampler_oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def rlwe_oracle():
    """oracle/rlwe_oracle.py (the example's RLWE flow over the C oracle)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("rlwe_oracle", os.path.join(ROOT, "oracle", "rlwe_oracle.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def rand_residues(rng, n, L, primes, bound_mult=1):
    """Seeded words in [0, bound_mult * P[r]) per residue r (numpy Generator)."""
    out = np.empty(n * L, dtype=np.uint64)
    for r in range(L):
        hi = int(primes[r]) * bound_mult
        out[r * n:(r + 1) * n] = rng.integers(0, hi, size=n, dtype=np.uint64)
    return out
