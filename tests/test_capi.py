"""CPU-side checks of the C ABI: the built library loads and exports every
entry point include/numa_hip.h declares (no device calls)."""
import ctypes
import os

import numpy as np
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
    assert ctypes.sizeof(H.KrylovParams) == 16 * 8
    assert ctypes.sizeof(H.KrylovResult) == 11 * 8
    assert ctypes.sizeof(H.GmresResult) == 6 * 8
    assert ctypes.sizeof(H.MGConfig) == 13 * 8
    assert ctypes.sizeof(H.SolverConfig) == 19 * 8
    assert ctypes.sizeof(H.NetTrainConfig) == 43 * 8
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


def test_uniform_host_is_mt19937_64():
    """std::mt19937_64 with the default seed 5489 gives 9981545732273789042 as
    its 10000th output (C++ standard, [rand.predef]); U = (u64 >> 11) 2^-53."""
    u = H.uniform_host(5489, 10000, 1.0)
    assert u[-1] == (9981545732273789042 >> 11) * 2.0 ** -53
    assert np.all((u >= 0.0) & (u < 1.0))
    v = H.uniform_host(5489, 10000, 0.1)
    np.testing.assert_array_equal(v, 0.1 * u)


def test_sweep_instances_are_seeded():
    n1, l1 = H.sweep_instances(20261203, 96)
    n2, l2 = H.sweep_instances(20261203, 96)
    np.testing.assert_array_equal(n1, n2)
    np.testing.assert_array_equal(l1, l2)
    assert set(n1.tolist()) <= {32, 48, 64, 96} and len(set(n1.tolist())) == 4
    assert l1.min() >= 1.0 and l1.max() <= 6.7
    # first draw: size index = u64 % 4 of the first output of mt19937_64(seed)
    n3, _ = H.sweep_instances(5489, 1, sizes=(10, 11, 12, 13))
    assert n3[0] == (10, 11, 12, 13)[14514284786278117030 % 4]
