"""The RLWE flow of the reference's example.cpp:12-58 restated over the
oracle's polynomial functions -- TEST INFRASTRUCTURE ONLY.

Only tests/, the package's smoke check and the bench's CPU baseline may use
this module, and only as a checker.

`orc` is the oracle wrapper of tests/_oracle.py (the C restatement in
oracle/yell_oracle.c: poly_forward / poly_forward_lazy / poly_backward,
details/poly.hpp:171-186; poly_add_product_of, :119-132; poly_negate,
:152-158 (0 -> p); poly_add, ops.hpp:26-41).  The samples are inputs: sx
(hwt), e0, e1, e2 (Gaussian), pk_ax and msg (uniform), u (ZO).

Pinned by: the same flow through the reference's own sources (yref_rlwe in
oracle/_ref, tests/test_rlwe.py), and by decryption: plain - msg is the
scheme's noise e1 - u*e0 + sx*e2 (negacyclic products), which the tests
recompute by schoolbook convolution.
"""


def keygen(orc, n, L, sx, e0, pk_ax):
    sx = orc.poly_forward(sx, n, L)                      # example.cpp:15
    bx = orc.poly_forward(e0, n, L)                      # :24-25
    bx = orc.poly_add_product_of(bx, pk_ax, sx, n, L)    # :26
    bx = orc.poly_negate(bx, n, L)                       # :27
    return sx, bx


def encrypt(orc, n, L, pk_bx, pk_ax, msg, u, e1, e2):
    u = orc.poly_forward(u, n, L)                        # :39-40
    bx = orc.poly_add(e1, msg, n, L)                     # :41-42
    bx = orc.poly_forward_lazy(bx, n, L)                 # :43
    bx = orc.poly_add_product_of(bx, u, pk_bx, n, L)     # :44
    ax = orc.poly_forward_lazy(e2, n, L)                 # :46-47
    ax = orc.poly_add_product_of(ax, u, pk_ax, n, L)     # :48
    return bx, ax


def decrypt(orc, n, L, sx, ct_bx, ct_ax):
    plain = orc.poly_add_product_of(ct_bx, sx, ct_ax, n, L)  # :55-56
    return orc.poly_backward(plain, n, L)                    # :57


def flow(orc, n, L, sx, e0, pk_ax, msg, u, e1, e2):
    """(sx in NTT form, pk_bx, ct_bx, ct_ax, plain)."""
    sxf, pk_bx = keygen(orc, n, L, sx, e0, pk_ax)
    ct_bx, ct_ax = encrypt(orc, n, L, pk_bx, pk_ax, msg, u, e1, e2)
    return sxf, pk_bx, ct_bx, ct_ax, decrypt(orc, n, L, sxf, ct_bx, ct_ax)
