"""The reference's own caller, example.cpp (toy RLWE: FastGaussianNoise,
gaussian<In, Out, D>, poly<N> samplers, transforms, operators,
add_product_of, negate, operator(), poly2mpz), compiles unchanged against the
drop-in headers under include/ and links against libyell_amd.so.  Read from
the reference checkout in place (nothing of it is kept here); skipped where
the checkout, GMP's header or the built library is absent."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = "/root/reference/example.cpp"
GMP_INC = "/opt/conda/include"
SYS_GMP = "/usr/lib/x86_64-linux-gnu/libgmp.so.10"
LIB_DIR = os.path.join(ROOT, "yell_amd", "lib")

needs_ref = pytest.mark.skipif(not os.path.exists(EXAMPLE) or not os.path.exists(os.path.join(GMP_INC, "gmp.h")),
                               reason="reference checkout or gmp.h absent")


def _gxx(*args):
    cmd = ["g++", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", GMP_INC] + list(args)
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


@needs_ref
def test_reference_example_compiles_unchanged():
    _gxx("-fsyntax-only", EXAMPLE)


@needs_ref
@pytest.mark.skipif(not os.path.exists(os.path.join(LIB_DIR, "libyell_amd.so")) or not os.path.exists(SYS_GMP),
                    reason="libyell_amd.so not built or system GMP absent")
def test_reference_example_links():
    with tempfile.TemporaryDirectory() as d:
        _gxx("-O1", EXAMPLE, "-o", os.path.join(d, "example"), "-L", LIB_DIR, "-lyell_amd", SYS_GMP)
