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
    assert ctypes.sizeof(H.MGConfig) == 13 * 8
    assert ctypes.sizeof(H.SolverConfig) == 19 * 8
    assert ctypes.sizeof(H.NetTrainConfig) == 40 * 8
    assert ctypes.sizeof(H.NetTrainStats) == 6 * 8


def test_pod_layouts_match_the_c_compiler(tmp_path):
    """sizeof / offsetof of every POD struct of include/numa_hip.h, from gcc,
    against the ctypes mirrors used by the Python bindings."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"numa_krylov_params": H.KrylovParams, "numa_krylov_result": H.KrylovResult,
               "numa_gmres_result": H.GmresResult, "numa_mg_config": H.MGConfig,
               "numa_solver_config": H.SolverConfig, "numa_net_train_config": H.NetTrainConfig,
               "numa_net_train_stats": H.NetTrainStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{H.INCLUDE}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        cname, what, val = line.split()
        got[(cname, what)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
