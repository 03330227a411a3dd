"""Runs the C++20 drop-in program (tests/cpp/test_dropin.cpp, built by
__graft_entry__.build()): the reference's own solver-path test cases written
against numa::hip."""
import os
import subprocess

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu


def test_cpp_dropin_program():
    exe = os.path.join(PKG, "build", "test_dropin")
    assert os.path.exists(exe), "build the package first (__graft_entry__.build())"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "dropin ok" in out.stdout
