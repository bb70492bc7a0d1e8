"""Runs the C++ drop-in test program (tests/cpp/test_shim.cpp) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "build", "test_shim")


def test_cpp_shim_program():
    assert os.path.exists(EXE), "build tests/cpp first (python3 yell_amd/build.py)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
