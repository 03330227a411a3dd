"""CPU-side checks of the C ABI: the built library loads and exports every
entry point include/numa_hip.h declares (no device calls)."""
import ctypes
import os

import pytest

import numa_hip as H


def test_header_declares_core_entry_points():
    syms = H.declared_symbols()
    for s in ("numa_hip_bratu_eval", "numa_hip_bratu_jvp_fd_mixed", "numa_hip_mg_apply",
              "numa_hip_gmres", "numa_hip_solve", "numa_hip_primitive_cost"):
        assert s in syms


@pytest.mark.skipif(not os.path.exists(H.LIB_PATH), reason="library not built")
def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(H.LIB_PATH)
    missing = [s for s in H.declared_symbols() if not hasattr(lib, s)]
    assert missing == []


def test_pod_layouts_are_8_byte_fields():
    assert ctypes.sizeof(H.KrylovParams) == 14 * 8
    assert ctypes.sizeof(H.KrylovResult) == 11 * 8
    assert ctypes.sizeof(H.GmresResult) == 6 * 8
    assert ctypes.sizeof(H.MGConfig) == 10 * 8
    assert ctypes.sizeof(H.SolverConfig) == 16 * 8
