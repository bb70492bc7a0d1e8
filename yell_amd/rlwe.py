"""The reference example's RLWE flow (example.cpp:12-58) on the device: key
generation, encryption and decryption of residue-major polynomials, each
step one C-ABI call on `stream`, the samples drawn by the device samplers.

The calls are the example's poly<N> calls in order (forward, forward_lazy,
add_product_of, negate, +=, backward), so the lazy ranges are the reference's:
the encryption's forward_lazy outputs in [0, 4p) go straight into
add_product_of, and the negated public key holds p where the product was 0.
Buffers are DeviceBuffer objects of degree * nmoduli words; every call is
stream-ordered and allocates nothing, so an encryption can be captured into a
graph (ya_graph_begin / ya_graph_end) once its calls have run once.

encrypt_many / decrypt_many (64-bit build) run `count` messages under one key
with one launch per step (the *_many transforms, samplers and pointwise
forms, the key a shared operand; the steps after each sampler fused into one
pass by default); message k gets the words encrypt / decrypt give it with
seeds + k.
"""
import numpy as np

from . import capi


class Rlwe:
    def __init__(self, degree, nmoduli, word=64, stream=None, hwt=64, sigma=3.2, rho=0x7F):
        self.n, self.L, self.word = degree, nmoduli, word
        self.pre = "ya_" if word == 64 else "ya32_"
        self.dt = np.uint64 if word == 64 else np.uint32
        self.stream = stream
        self.hwt, self.sigma, self.rho = hwt, sigma, rho
        self.nbytes = degree * nmoduli * np.dtype(self.dt).itemsize

    def buffer(self, count=1):
        return capi.DeviceBuffer(self.n * self.L * count, self.dt)

    def _op(self, fn, *bufs):
        capi.call(self.pre + fn, *[b.ptr for b in bufs], self.n, self.L, self.stream)

    def _sample(self, kind, buf, param, seed):
        capi.sample_device(kind, buf, self.n, self.L, param, seed, stream=self.stream, word=self.word)

    def keygen(self, sx, pk_bx, pk_ax, seeds):
        """SK and PK (example.cpp:12-31); seeds = (sx, e0, ax)."""
        self._sample("hwt", sx, self.hwt, seeds[0])          # :14
        self._op("poly_forward", sx)                          # :15
        self._sample("uniform", pk_ax, None, seeds[2])        # :23
        self._sample("gaussian", pk_bx, self.sigma, seeds[1])  # :24
        self._op("poly_forward", pk_bx)                       # :25
        self._op("poly_add_product_of", pk_bx, pk_ax, sx)     # :26
        self._op("poly_negate", pk_bx)                        # :27

    def encrypt(self, ct_bx, ct_ax, u, msg, pk_bx, pk_ax, seeds):
        """CTXT (example.cpp:33-50); seeds = (u, e1, e2); u is work space."""
        self._sample("zo", u, self.rho, seeds[0])             # :39
        self._op("poly_forward", u)                           # :40
        self._sample("gaussian", ct_bx, self.sigma, seeds[1])  # :41
        self._op("poly_add", ct_bx, ct_bx, msg)               # :42
        self._op("poly_forward_lazy", ct_bx)                  # :43
        self._op("poly_add_product_of", ct_bx, u, pk_bx)      # :44
        self._sample("gaussian", ct_ax, self.sigma, seeds[2])  # :46
        self._op("poly_forward_lazy", ct_ax)                  # :47
        self._op("poly_add_product_of", ct_ax, u, pk_ax)      # :48

    def decrypt(self, plain, ct_bx, ct_ax, sx):
        """decrypt (example.cpp:52-58)."""
        capi.call("ya_memcpy_dtod", plain.ptr, ct_bx.ptr, self.nbytes, self.stream)  # :55
        self._op("poly_add_product_of", plain, sx, ct_ax)     # :56
        self._op("poly_backward", plain)                      # :57

    # ---- many messages under one key, one launch per step (64-bit build) ----
    def _many(self, fn, *bufs, count, shared=None):
        extra = () if shared is None else (shared,)
        capi.call("ya_poly_" + fn, *[b.ptr for b in bufs], self.n, self.L, count, *extra, self.stream)

    def encrypt_many(self, ct_bx, ct_ax, u, msgs, pk_bx, pk_ax, count, seeds, fused=True):
        """encrypt() of messages 0..count-1 (polynomials back to back in msgs)
        with seeds (seeds[0] + k, seeds[1] + k, seeds[2] + k).  fused: each
        ciphertext half is one ya_poly_forward_lazy_add_product_many call
        (the addition of the message, forward_lazy and add_product_of in one
        pass); else the separate launches.  The same words either way."""
        assert self.word == 64
        n, L, st = self.n, self.L, self.stream
        capi.call("ya_sample_zo_device_many", u.ptr, n, L, count, self.rho, seeds[0], st)
        self._many("forward_many", u, count=count)
        capi.call("ya_sample_gaussian_device_many", ct_bx.ptr, n, L, count, self.sigma, 1, seeds[1], st)
        if fused:
            capi.call("ya_poly_forward_lazy_add_product_many", ct_bx.ptr, msgs.ptr, u.ptr, pk_bx.ptr, n, L, count, st)
        else:
            self._many("add_many", ct_bx, ct_bx, msgs, count=count, shared=0)
            self._many("forward_lazy_many", ct_bx, count=count)
            self._many("add_product_of_many", ct_bx, u, pk_bx, count=count, shared=1)
        capi.call("ya_sample_gaussian_device_many", ct_ax.ptr, n, L, count, self.sigma, 1, seeds[2], st)
        if fused:
            capi.call("ya_poly_forward_lazy_add_product_many", ct_ax.ptr, None, u.ptr, pk_ax.ptr, n, L, count, st)
        else:
            self._many("forward_lazy_many", ct_ax, count=count)
            self._many("add_product_of_many", ct_ax, u, pk_ax, count=count, shared=1)

    def decrypt_many(self, plain, ct_bx, ct_ax, sx, count, fused=True):
        """decrypt() of ciphertexts 0..count-1; sx * ct_ax is taken as
        ct_ax * sx (the Barrett muladd of x * y + acc is symmetric in x, y).
        fused: one ya_poly_add_product_backward_many call; else a copy and
        the separate launches."""
        assert self.word == 64
        n, L, st = self.n, self.L, self.stream
        if fused:
            capi.call("ya_poly_add_product_backward_many", plain.ptr, ct_bx.ptr, ct_ax.ptr, sx.ptr, n, L, count, st)
            return
        capi.call("ya_memcpy_dtod", plain.ptr, ct_bx.ptr, self.nbytes * count, st)
        self._many("add_product_of_many", plain, ct_ax, sx, count=count, shared=1)
        self._many("backward_many", plain, count=count)
