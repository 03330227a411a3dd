"""The no-allocation FD J v entry point (numa_hip_bratu_jvp_fd_noalloc), called
through the C ABI, against the context entry point numa_hip_bratu_jvp_fd on
the same inputs: bit for bit, on two calls in a row (the second on a smaller
grid, so the shared scratch array keeps the first call's contents), then
against the oracle's forward difference."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel
import numa_hip as H

pytestmark = pytest.mark.gpu


def _hip_device_synchronize() -> None:
    # the no-allocation call runs on the null stream; the context's stream is
    # non-blocking, so wait for the device before reading through the context
    hip = C.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize.restype = C.c_int
    assert hip.hipDeviceSynchronize() == 0


def _inputs(ctx, n, seed):
    rng = np.random.default_rng(seed)
    x = 0.1 + 0.05 * rng.standard_normal(n ** 3)
    v = rng.standard_normal(n ** 3)
    op = H.Bratu3D(ctx, n, 6.0)
    dx, dv, dF = ctx.array(x), ctx.array(v), ctx.empty(n ** 3)
    op.eval(dx, dF)
    return op, x, v, dx, dv, dF


def test_noalloc_jvp_fd_matches_context_entry(ctx, oracle):
    lib = H.load()
    for n, seed in ((40, 1), (24, 2)):  # second call smaller
        print(f"noalloc jvp_fd: n={n} seed={seed}", flush=True)
        op, x, v, dx, dv, dF = _inputs(ctx, n, seed)
        ref, got = ctx.empty(n ** 3), ctx.empty(n ** 3)
        op.jvp_fd(dx, dF, dv, ref)
        ctx.sync()
        H.check(lib.numa_hip_bratu_jvp_fd_noalloc(0, None, n, 6.0, dx.ptr, dF.ptr, dv.ptr,
                                                  got.ptr))
        _hip_device_synchronize()
        a, b = ref.numpy(), got.numpy()
        np.testing.assert_array_equal(b, a)
        want = oracle.bratu_jvp_fd(n, 6.0, x, v)
        assert rel(b, want) < 1e-7, rel(b, want)
        for d in (dx, dv, dF, ref, got):
            d.free()


def test_noalloc_rejects_bad_arguments(ctx):
    lib = H.load()
    d = ctx.empty(8)
    assert lib.numa_hip_bratu_jvp_fd_noalloc(0, None, 4096, 1.0, d.ptr, d.ptr, d.ptr,
                                             d.ptr) == 4  # NUMA_ERR_OVERFLOW
    assert lib.numa_hip_bratu_jvp_fd_noalloc(0, None, 2, 1.0, None, d.ptr, d.ptr,
                                             d.ptr) == 1  # NUMA_ERR_ARG
    assert lib.numa_hip_bratu_jvp_fd_noalloc(-1, None, 2, 1.0, d.ptr, d.ptr, d.ptr,
                                             d.ptr) == 1  # NUMA_ERR_ARG
    d.free()
