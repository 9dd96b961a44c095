"""Runs the C++ program that calls the gtom:: drop-in header (tests/cpp/test_shim.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "test_shim")


def test_cpp_dropin_program():
    assert os.path.exists(BIN), "build the C++ test with __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all passed" in r.stdout
