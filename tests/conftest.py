"""Shared test setup: import paths for the oracle (test infrastructure) and the
HIP library's Python bindings, and the `gpu` marker."""
import os
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "chebyshev-nonlinear-lab_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)

# autotuner cache of this test session (read by the library at its first use)
os.environ.setdefault("NUMA_HIP_TUNE_CACHE",
                      os.path.join(tempfile.gettempdir(), f"numa_hip_tune_test_{os.getpid()}.json"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    if not os.path.exists(O.ORACLE_SO):
        O.build()
    return O.Oracle()


@pytest.fixture(scope="session")
def ctx():
    import numa_hip as H
    c = H.Context(0)
    yield c
    c.close()


def rel(a, b):
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))
