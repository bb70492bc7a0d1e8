"""Runs the C++ drop-in test program (tests/cpp/test_shim.cpp) on the GPU, as
built for the 64-bit moduli (test_shim) and with -DYELL_USE_32BITS_MODULI
(test_shim32)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "build")


@pytest.mark.parametrize("name", ["test_shim", "test_shim32"])
def test_cpp_shim_program(name):
    exe = os.path.join(BUILD, name)
    assert os.path.exists(exe), "build tests/cpp first (python3 yell_amd/build.py)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
