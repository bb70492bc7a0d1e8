"""yell_amd: MI355X (gfx950) engine for YELL's negacyclic NTT / RNS polynomial path.

The product is the C-ABI shared library yell_amd/lib/libyell_amd.so
(include/yell_amd.h) and the C++ drop-in headers under include/yell/.  This
Python package only binds the C ABI (yell_amd.capi) for tests and the bench.
"""
from . import capi  # noqa: F401
