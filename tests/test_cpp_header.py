"""CPU check that the C++20 drop-in headers compile on their own and satisfy
the reference's operator concepts (static_asserts in test_dropin.cpp)."""
import os
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_dropin_headers_compile():
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Werror",
                              "-I" + os.path.join(ROOT, "include"),
                              os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp")],
                             capture_output=True, text=True, cwd=d)
        assert out.returncode == 0, out.stderr
