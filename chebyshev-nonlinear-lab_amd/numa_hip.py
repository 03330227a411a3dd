"""ctypes bindings of lib/libnuma_hip.so (include/numa_hip.h).

This is the Python face of the MI355X backend used by the tests and bench.py.
It mirrors the reference's interface for the solver path
(include/numa/methods/krylov_newton.hpp, linalg/gmres.hpp,
linalg/poisson_mg.hpp, problems/bratu3d.hpp): KrylovParams / KrylovResult /
GmresResult with the same fields, Bratu3D with eval / jvp / d2, PoissonMG
with levels() and operator(), newton_krylov / chebyshev_inexact / gmres.
Device memory is owned by DeviceArray (allocated through the C ABI, no torch).

There is no CPU fallback: loading fails loudly if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnuma_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include", "numa_hip.h")

_i64 = C.c_int64
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)

INJECT_F, INJECT_JV = 0, 1
NUMA_OK = 0
ERR_NAMES = {1: "invalid argument", 2: "HIP error", 3: "out of memory", 4: "overflow",
             5: "bad state"}

METHOD_NEWTON, METHOD_CHEBYSHEV, METHOD_TF2018 = 0, 1, 2
PRECOND_NONE, PRECOND_JACOBI, PRECOND_GMG, PRECOND_LEARNED = 0, 1, 2, 3
PRECISION_FP64, PRECISION_MIXED = 0, 1
DERIV_FD, DERIV_ANALYTIC = 0, 1
FORCING_EW, FORCING_CONSTANT = 0, 1
SMOOTH_RBGS, SMOOTH_WJACOBI, SMOOTH_CHEBYSHEV = 0, 1, 2
FUSE_NONE, FUSE_SWEEP, FUSE_PHASE, FUSE_AUTO = 0, 1, 2, 3
# MGConfig.rows < 0: box form of fused smoothing (at most 3 steps per launch),
# output box i x j x k per workgroup
BOX_8x8x16, BOX_4x8x16, BOX_4x4x16 = -1, -2, -3
BOX_EDGES = {BOX_8x8x16: (8, 8, 16), BOX_4x8x16: (4, 8, 16), BOX_4x4x16: (4, 4, 16)}
PRIM_F, PRIM_JVP, PRIM_D2, PRIM_PREC, PRIM_SMOOTH, PRIM_MGS, PRIM_COPY = 0, 1, 2, 3, 4, 5, 6
LOSS_CORRECTION, LOSS_RESIDUAL = 0, 1


class KrylovParams(C.Structure):
    _fields_ = [("rtol_nl", C.c_double), ("max_nl", _i64), ("eta_init", C.c_double),
                ("eta_max", C.c_double), ("gamma", C.c_double), ("alpha", C.c_double),
                ("eta2_factor", C.c_double), ("eta2_max", C.c_double),
                ("gmres_restart", _i64), ("gmres_max_iters", _i64), ("armijo", C.c_double),
                ("min_lambda", C.c_double), ("recycle", _i64), ("forcing", _i64),
                ("tight_res", C.c_double), ("tight_eta", C.c_double)]


class KrylovResult(C.Structure):
    _fields_ = [("converged", _i64), ("iters", _i64), ("n_F", _i64), ("n_JVP", _i64),
                ("n_D2", _i64), ("n_lin_iters", _i64), ("n_lin_failures", _i64),
                ("n_prec", _i64), ("res0", C.c_double), ("res", C.c_double),
                ("t_total", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(C.Structure):
    """numa_batch_stats: counters of a batched solver."""
    _fields_ = [("launches", _i64), ("graph_launches", _i64), ("arnoldi_steps", _i64),
                ("host_syncs", _i64), ("bytes", C.c_double), ("tables", _i64),
                ("table_bytes", _i64), ("graphs", _i64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class GmresResult(C.Structure):
    _fields_ = [("iters", _i64), ("matvecs", _i64), ("precs", _i64), ("rel_res", C.c_double),
                ("converged", _i64), ("breakdown", _i64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MGConfig(C.Structure):
    _fields_ = [("smoother", _i64), ("pre", _i64), ("post", _i64), ("degree", _i64),
                ("max_levels", _i64), ("min_coarse_n", _i64), ("coarse_iters", _i64),
                ("even_coarsening", _i64), ("omega", C.c_double), ("cheb_ratio", C.c_double),
                ("fuse", _i64), ("ichunk", _i64), ("rows", _i64)]


class SolverConfig(C.Structure):
    _fields_ = [("method", _i64), ("precond", _i64), ("precision", _i64), ("flexible", _i64),
                ("jv_mode", _i64), ("d2_mode", _i64), ("mg", MGConfig)]


class EigenResult(C.Structure):
    """numa::analysis::EigenEstimate (analysis/spectral.hpp:22-32) without the vector."""
    _fields_ = [("converged", _i64), ("eigenvalue", C.c_double), ("residual_rel", C.c_double),
                ("outer_iterations", _i64), ("linear_iterations", _i64),
                ("linear_matvecs", _i64), ("preconditioner_applications", _i64),
                ("linear_failures", _i64)]


class EigenParams(C.Structure):
    _fields_ = [("eigen_rtol", C.c_double), ("linear_rtol", C.c_double), ("max_outer", _i64),
                ("restart", _i64), ("max_linear", _i64)]


class ContinuationParams(C.Structure):
    _fields_ = [("lambda_start", C.c_double), ("lambda_stop", C.c_double), ("step", C.c_double),
                ("min_step", C.c_double), ("spectral_stop", C.c_double),
                ("eigen_rtol", C.c_double), ("linear_rtol", C.c_double)]


class ContinuationPoint(C.Structure):
    """One CSV row of apps/calibrate_b6.cpp:128-137."""
    _fields_ = [("lambda_", C.c_double), ("spectral_gap", C.c_double),
                ("eigen_residual", C.c_double), ("continuation_iters", _i64), ("n_F", _i64),
                ("n_JVP", _i64), ("n_prec", _i64)]


def eigen_params(eigen_rtol=1e-7, linear_rtol=1e-10, max_outer=30, restart=100,
                 max_linear=2000) -> EigenParams:
    """inverse_smallest_eigenpair defaults (analysis/spectral.hpp:34-42)."""
    return EigenParams(eigen_rtol, linear_rtol, max_outer, restart, max_linear)


def continuation_params(**kw) -> ContinuationParams:
    """calibrate_b6 defaults (apps/calibrate_b6.cpp:30-40)."""
    d = dict(lambda_start=0.0, lambda_stop=9.8, step=0.25, min_step=1e-4, spectral_stop=5e-2,
             eigen_rtol=1e-7, linear_rtol=1e-10)
    d.update(kw)
    return ContinuationParams(**d)


def corrector_params() -> "KrylovParams":
    """Newton-Krylov parameters of the continuation corrector (calibrate_b6.cpp:181-187)."""
    return krylov_params(rtol_nl=1e-11, max_nl=100, eta_init=1e-8, eta_max=1e-6,
                         gmres_restart=100, gmres_max_iters=2000)


FOLD_CSV_HEADER = "lambda,spectral_gap,eigen_residual,continuation_iters,n_F,n_JVP,n_prec"


def fold_csv_line(p: "ContinuationPoint") -> str:
    """A calibrate_b6 CSV row (calibrate_b6.cpp:128-137: default stream format,
    6 significant digits)."""
    return "%g,%g,%g,%d,%d,%d,%d" % (p.lambda_, p.spectral_gap, p.eigen_residual,
                                     p.continuation_iters, p.n_F, p.n_JVP, p.n_prec)


class Timers(C.Structure):
    """numa::op::Timers (operators/timing.hpp:16-24): seconds per primitive."""
    _fields_ = [("t_F", C.c_double), ("t_JVP", C.c_double), ("t_D2", C.c_double),
                ("t_prec", C.c_double), ("t_setup_dynamic", C.c_double)]


class NetTrainConfig(C.Structure):
    _fields_ = [("pairs", _i64), ("batch", _i64), ("sizes", _i64 * 4), ("lambda_lo", C.c_double),
                ("lambda_hi", C.c_double), ("noise", C.c_double), ("lr", C.c_double),
                ("seed", _i64), ("gmres_restart", _i64), ("mg", MGConfig),
                ("solve", KrylovParams), ("loss", _i64), ("lr_end", C.c_double)]


class NetTrainStats(C.Structure):
    _fields_ = [("pairs", _i64), ("steps", _i64), ("solves", _i64), ("loss_first", C.c_double),
                ("loss_last", C.c_double), ("seconds", C.c_double)]


def krylov_params(**kw) -> KrylovParams:
    """Defaults of numa::op::KrylovParams (methods/krylov_newton.hpp:24-38)."""
    p = KrylovParams(rtol_nl=1e-8, max_nl=50, eta_init=0.5, eta_max=0.9, gamma=0.9, alpha=2.0,
                     eta2_factor=10.0, eta2_max=0.5, gmres_restart=50, gmres_max_iters=500,
                     armijo=1e-4, min_lambda=1e-4, recycle=0, forcing=FORCING_EW,
                     tight_res=0.0, tight_eta=0.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def mg_reference() -> MGConfig:
    """PoissonMG(n) (linalg/poisson_mg.hpp:26-64): RB-GS V(2,2), odd-n coarsening."""
    return MGConfig(smoother=SMOOTH_RBGS, pre=2, post=2, degree=3, max_levels=0, min_coarse_n=1,
                    coarse_iters=30, even_coarsening=0, omega=6.0 / 7.0, cheb_ratio=4.0)


def mg_config(**kw) -> MGConfig:
    m = mg_reference()
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def mg_cheb(levels: int = 5, degree: int = 3, pre: int = 2, post: int = 2,
            cheb_ratio: float = 4.0, coarse_iters: int = 24, fuse: int = FUSE_NONE,
            ichunk: int = 0, rows: int = 0) -> MGConfig:
    """5-level even-n V(2,2) with the degree-3 Chebyshev smoother (256^3 config)."""
    return mg_config(smoother=SMOOTH_CHEBYSHEV, pre=pre, post=post, degree=degree,
                     max_levels=levels, min_coarse_n=2, coarse_iters=coarse_iters,
                     even_coarsening=1, cheb_ratio=cheb_ratio, fuse=fuse, ichunk=ichunk,
                     rows=rows)


def mg_wjacobi(levels: int = 0, pre: int = 2, post: int = 2, omega: float = 6.0 / 7.0,
               coarse_iters: int = 24, fuse: int = FUSE_NONE, ichunk: int = 0,
               rows: int = 0) -> MGConfig:
    """Even-n V(2,2) with the weighted-Jacobi smoother (32^3 CPU config)."""
    return mg_config(smoother=SMOOTH_WJACOBI, pre=pre, post=post, max_levels=levels,
                     min_coarse_n=2, coarse_iters=coarse_iters, even_coarsening=1, omega=omega,
                     fuse=fuse, ichunk=ichunk, rows=rows)


def solver_config(method=METHOD_NEWTON, precond=PRECOND_GMG, precision=PRECISION_FP64,
                  flexible=0, jv_mode=DERIV_FD, d2_mode=DERIV_FD, mg=None) -> SolverConfig:
    return SolverConfig(method=method, precond=precond, precision=precision, flexible=flexible,
                        jv_mode=jv_mode, d2_mode=d2_mode,
                        mg=mg if mg is not None else mg_reference())


class NumaError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NumaError(f"{path} not found: build it with __graft_entry__.build() "
                        "(there is no CPU fallback)")
    L = C.CDLL(path)
    L.numa_hip_last_error.restype = C.c_char_p
    sig = {
        "numa_hip_ctx_create": [C.c_int, C.POINTER(_vp)],
        "numa_hip_ctx_destroy": [_vp],
        "numa_hip_ctx_sync": [_vp],
        "numa_hip_malloc": [_vp, C.c_size_t, C.POINTER(_vp)],
        "numa_hip_free": [_vp, _vp],
        "numa_hip_memcpy_h2d": [_vp, _vp, _vp, C.c_size_t],
        "numa_hip_memcpy_d2h": [_vp, _vp, _vp, C.c_size_t],
        "numa_hip_bratu_eval": [_vp, _i64, C.c_double, _vp, _vp, _dp],
        "numa_hip_bratu_jvp": [_vp, _i64, C.c_double, _vp, _vp, _vp],
        "numa_hip_bratu_d2": [_vp, _i64, C.c_double, _vp, _vp, _vp],
        "numa_hip_bratu_jacobi_diag": [_vp, _i64, C.c_double, _vp, _vp],
        "numa_hip_bratu_jvp_fd": [_vp, _i64, C.c_double, _vp, _vp, _vp, _vp],
        "numa_hip_bratu_jvp_fd_mixed": [_vp, _i64, C.c_double, _vp, _vp, _vp, _vp],
        "numa_hip_bratu_jvp_fd_rearranged": [_vp, _i64, C.c_double, _vp, _vp, _vp, _vp],
        "numa_hip_bratu_d2_fd": [_vp, _i64, C.c_double, _vp, _vp, _vp, _vp],
        "numa_hip_bratu_jvp_fd_noalloc": [C.c_int, _vp, _i64, C.c_double, _vp, _vp, _vp, _vp],
        "numa_hip_dot": [_vp, _i64, _vp, _vp, _dp],
        "numa_hip_dot_f32": [_vp, _i64, _vp, _vp, _dp],
        "numa_hip_axpy": [_vp, _i64, C.c_double, _vp, _vp],
        "numa_hip_mg_create": [_vp, _i64, C.POINTER(MGConfig), _i64, C.POINTER(_vp)],
        "numa_hip_mg_levels": [_vp, C.POINTER(_i64), C.POINTER(_i64)],
        "numa_hip_mg_apply": [_vp, _vp, _vp],
        "numa_hip_mg_apply_norm": [_vp, _vp, _vp, C.POINTER(C.c_double)],
        "numa_hip_mg_destroy": [_vp],
        "numa_hip_bratu_solver_create": [_vp, _i64, C.c_double, C.POINTER(SolverConfig), _i64,
                                         C.POINTER(_vp)],
        "numa_hip_solver_destroy": [_vp],
        "numa_hip_gmres": [_vp, _vp, _vp, _vp, C.c_double, _i64, _i64, _i64,
                           C.POINTER(GmresResult)],
        "numa_hip_solve": [_vp, C.POINTER(KrylovParams), _vp, C.POINTER(KrylovResult)],
        "numa_hip_solve_host": [_vp, C.POINTER(KrylovParams), _dp, C.POINTER(KrylovResult)],
        "numa_hip_primitive_cost": [_vp, _i64, _i64, _dp, _dp],
        "numa_hip_mfma_peak": [_vp, _i64, _dp],
        "numa_hip_chandra_create": [_vp, _i64, _i64, _dp, _i64, _i64, _i64, _i64, C.POINTER(_vp)],
        "numa_hip_chandra_destroy": [_vp],
        "numa_hip_chandra_eval": [_vp, _vp, _vp, _dp],
        "numa_hip_chandra_jvp": [_vp, _vp, _vp, _vp],
        "numa_hip_chandra_solve": [_vp, C.POINTER(KrylovParams), _vp, C.POINTER(KrylovResult)],
        "numa_hip_chandra_gemm_cost": [_vp, _i64, _i64, _dp, _dp],
        "numa_hip_chandra_d2": [_vp, _vp, _vp, _vp],
        "numa_hip_chandra_set_method": [_vp, _i64],
        "numa_hip_chandra_quadrature": [_i64, _i64, _dp, _dp],
        "numa_hip_net_create": [_vp, _i64, C.POINTER(_vp)],
        "numa_hip_net_destroy": [_vp],
        "numa_hip_net_param_count": [C.POINTER(_i64)],
        "numa_hip_net_get_params": [_vp, _vp],
        "numa_hip_net_set_params": [_vp, _vp],
        "numa_hip_net_forward": [_vp, _i64, _i64, _vp, _vp],
        "numa_hip_net_loss_grad": [_vp, _i64, _i64, _vp, _vp, _i64, _dp, _vp],
        "numa_hip_net_train_step": [_vp, _i64, _i64, _vp, _vp, _i64, C.c_double, _dp],
        "numa_hip_net_apply": [_vp, _i64, _vp, _vp],
        "numa_hip_net_train_solver": [_vp, C.POINTER(NetTrainConfig), _dp, _i64,
                                      C.POINTER(NetTrainStats)],
        "numa_hip_net_cost": [_vp, _i64, _i64, _i64, _i64, _dp],
        "numa_hip_solver_set_net": [_vp, _vp, C.c_double],
        "numa_hip_solver_learned_stats": [_vp, C.POINTER(_i64)],
        "numa_hip_solver_inject_nan": [_vp, _i64, _i64],
        "numa_hip_mg_fusion": [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)],
        "numa_hip_mg_set_resid_restrict": [_vp, _i64],
        "numa_hip_solver_set_graphs": [_vp, _i64],
        "numa_hip_tf2018_curvature": [_vp, _vp, _vp, _vp, C.POINTER(C.c_double)],
        "numa_hip_eigen_smallest": [_vp, _vp, C.POINTER(EigenParams), C.POINTER(EigenResult),
                                    _vp],
        "numa_hip_start_write": [C.c_char_p, _vp, _i64],
        "numa_hip_start_read": [C.c_char_p, _i64, _vp],
        "numa_hip_sha256": [_vp, _i64, C.c_char_p],
        "numa_hip_sha256_file": [C.c_char_p, C.c_char_p],
        "numa_hip_perturb_host": [_vp, _i64, C.c_double, C.c_uint64, _vp],
        "numa_hip_prolong3d": [_vp, _vp, _i64, _vp],
        "numa_hip_continuation": [_vp, _i64, C.POINTER(SolverConfig),
                                  C.POINTER(ContinuationParams), C.POINTER(KrylovParams),
                                  C.POINTER(ContinuationPoint), _i64, C.POINTER(_i64),
                                  C.POINTER(_i64), C.POINTER(C.c_double), _vp],
        "numa_hip_solver_set_timing": [_vp, _i64],
        "numa_hip_solver_timers": [_vp, C.POINTER(Timers), _i64],
        "numa_hip_solver_graphs_cached": [_vp, C.POINTER(_i64)],
        "numa_hip_sweep_create": [_vp, _i64, C.POINTER(SolverConfig), _i64, C.POINTER(_vp)],
        "numa_hip_sweep_destroy": [_vp],
        "numa_hip_sweep_run": [_vp, _i64, C.POINTER(_i64), _dp, C.POINTER(_vp),
                               C.POINTER(KrylovParams), _i64, C.POINTER(KrylovResult), _dp],
        "numa_hip_sweep_instances": [_i64, _i64, C.POINTER(_i64), _i64, C.c_double, C.c_double,
                                     C.POINTER(_i64), _dp],
        "numa_hip_uniform_host": [_i64, _i64, C.c_double, _dp],
        "numa_hip_batch_create": [_vp, _i64, C.POINTER(_i64), _dp, C.POINTER(SolverConfig), _i64,
                                  C.POINTER(_vp)],
        "numa_hip_batch_destroy": [_vp],
        "numa_hip_batch_set_graphs": [_vp, _i64],
        "numa_hip_solver_sweep_events": [_vp, _i64],
        "numa_hip_solver_sweep_time": [_vp, _dp, C.POINTER(_i64)],
        "numa_hip_batch_set_smoothing": [_vp, _i64, C.POINTER(_i64)],
        "numa_hip_batch_solve": [_vp, C.POINTER(KrylovParams), C.POINTER(_vp),
                                 C.POINTER(KrylovResult), _dp],
        "numa_hip_batch_stats": [_vp, _i64, C.POINTER(BatchStats)],
        "numa_hip_solver_mg_fusion": [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64),
                                      C.POINTER(_i64)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != NUMA_OK:
        msg = _lib.numa_hip_last_error().decode() if _lib is not None else ""
        raise NumaError(f"numa_hip error {rc} ({ERR_NAMES.get(rc, '?')}): {msg}")


def declared_symbols(header: str = INCLUDE) -> list[str]:
    """Entry points declared in include/numa_hip.h."""
    import re
    txt = open(header).read()
    return sorted(set(re.findall(r"\b(numa_hip_[a-z0-9_]+)\s*\(", txt)))


class Context:
    """A device + stream (numa_hip_ctx)."""

    def __init__(self, device: int = 0):
        import weakref
        self.lib = load()
        h = _vp()
        check(self.lib.numa_hip_ctx_create(device, C.byref(h)))
        self.h = h
        # handles created on this context; closed before the context itself
        # (a context must outlive its solvers, multigrids and buffers)
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)
        return obj

    def sync(self):
        check(self.lib.numa_hip_ctx_sync(self.h))

    def mfma_peak(self, kind: int) -> float:
        """Measured dense MFMA TFLOP/s: kind 0 bf16 (32x32x16), 1 fp64 (16x16x4)."""
        out = C.c_double()
        check(self.lib.numa_hip_mfma_peak(self.h, kind, C.byref(out)))
        return out.value

    def close(self):
        if self.h:
            for child in list(self._children):
                child.close()
            check(self.lib.numa_hip_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, a) -> "DeviceArray":
        return DeviceArray.from_numpy(self, np.asarray(a))

    def empty(self, n: int, dtype=np.float64) -> "DeviceArray":
        return DeviceArray(self, n, dtype)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceArray:
    """A 1-D device buffer of float64 or float32 owned through the C ABI."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        p = _vp()
        check(ctx.lib.numa_hip_malloc(ctx.h, max(1, self.nbytes), C.byref(p)))
        self.ptr = p
        ctx._adopt(self)

    @property
    def nbytes(self) -> int:
        return self.n * self.dtype.itemsize

    @classmethod
    def from_numpy(cls, ctx: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        assert a.dtype in (np.float64, np.float32)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        check(self.ctx.lib.numa_hip_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        check(self.ctx.lib.numa_hip_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            check(self.ctx.lib.numa_hip_free(self.ctx.h, self.ptr))
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Bratu3D:
    """problems/bratu3d.hpp:12-100 on the device: F(u) = -Lap_h u - lambda e^u."""

    def __init__(self, ctx: Context, n: int, lam: float = 1.0):
        self.ctx, self.n, self.lam = ctx, int(n), float(lam)

    def size(self) -> int:
        return self.n ** 3

    def eval(self, u: DeviceArray, out: DeviceArray) -> float:
        nrm = C.c_double()
        check(self.ctx.lib.numa_hip_bratu_eval(self.ctx.h, self.n, self.lam, u.ptr, out.ptr,
                                               C.byref(nrm)))
        return nrm.value

    def jvp(self, u, v, out):
        check(self.ctx.lib.numa_hip_bratu_jvp(self.ctx.h, self.n, self.lam, u.ptr, v.ptr, out.ptr))

    def d2(self, u, v, out):
        check(self.ctx.lib.numa_hip_bratu_d2(self.ctx.h, self.n, self.lam, u.ptr, v.ptr, out.ptr))

    def jacobi_diag(self, u, d):
        check(self.ctx.lib.numa_hip_bratu_jacobi_diag(self.ctx.h, self.n, self.lam, u.ptr, d.ptr))

    def jvp_fd(self, x, F0, v, out):
        f = (self.ctx.lib.numa_hip_bratu_jvp_fd_mixed if v.dtype == np.float32
             else self.ctx.lib.numa_hip_bratu_jvp_fd)
        check(f(self.ctx.h, self.n, self.lam, x.ptr, F0.ptr, v.ptr, out.ptr))

    def jvp_fd_rearranged(self, x, v, E, out):
        """The mixed solver's J v: x FP64; v, E (= lambda e^x, written) and
        out FP32."""
        check(self.ctx.lib.numa_hip_bratu_jvp_fd_rearranged(self.ctx.h, self.n, self.lam, x.ptr,
                                                            v.ptr, E.ptr, out.ptr))

    def d2_fd(self, x, F0, v, out):
        check(self.ctx.lib.numa_hip_bratu_d2_fd(self.ctx.h, self.n, self.lam, x.ptr, F0.ptr,
                                                v.ptr, out.ptr))


class PoissonMG:
    """linalg/poisson_mg.hpp:24-299 on the device (plus even-n / WJ / Chebyshev)."""

    def __init__(self, ctx: Context, n: int, cfg: MGConfig | None = None,
                 precision: int = PRECISION_FP64):
        self.ctx, self.n, self.precision = ctx, int(n), precision
        self.cfg = cfg if cfg is not None else mg_reference()
        h = _vp()
        check(ctx.lib.numa_hip_mg_create(ctx.h, self.n, C.byref(self.cfg), precision, C.byref(h)))
        self.h = h
        ctx._adopt(self)

    def levels(self) -> int:
        L = _i64()
        check(self.ctx.lib.numa_hip_mg_levels(self.h, C.byref(L), None))
        return L.value

    def sizes(self) -> list[int]:
        L = _i64()
        s = (_i64 * 64)()
        check(self.ctx.lib.numa_hip_mg_levels(self.h, C.byref(L), s))
        return [int(s[i]) for i in range(L.value)]

    def __call__(self, r: DeviceArray, z: DeviceArray) -> None:
        check(self.ctx.lib.numa_hip_mg_apply(self.h, r.ptr, z.ptr))

    def apply_norm(self, r: DeviceArray, z: DeviceArray) -> float:
        """z = V-cycle(r) and ||z||_2 (from the last launch when it is fused)."""
        out = C.c_double()
        check(self.ctx.lib.numa_hip_mg_apply_norm(self.h, r.ptr, z.ptr, C.byref(out)))
        return out.value

    def fusion(self) -> list[tuple[int, int, int]]:
        """(NUMA_FUSE_* mode, plane chunk, rows per thread) per level."""
        L = self.levels()
        m, ic, rw = (_i64 * L)(), (_i64 * L)(), (_i64 * L)()
        check(self.ctx.lib.numa_hip_mg_fusion(self.h, m, ic, rw))
        return [(int(m[i]), int(ic[i]), int(rw[i])) for i in range(L)]

    def set_resid_restrict(self, ich: int) -> None:
        """Residual + restriction as one launch with `ich` coarse planes per
        workgroup (> 0), the default chunk (< 0) or two launches (0)."""
        check(self.ctx.lib.numa_hip_mg_set_resid_restrict(self.h, int(ich)))

    def close(self):
        if self.h and self.ctx.h:
            check(self.ctx.lib.numa_hip_mg_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BratuSolver:
    """Device-resident Newton-Krylov / inexact Chebyshev on Bratu3D."""

    def __init__(self, ctx: Context, n: int, lam: float, cfg: SolverConfig, restart: int = 50):
        self.ctx, self.n, self.lam, self.cfg = ctx, int(n), float(lam), cfg
        h = _vp()
        check(ctx.lib.numa_hip_bratu_solver_create(ctx.h, self.n, self.lam, C.byref(cfg),
                                                   int(restart), C.byref(h)))
        self.h = h
        ctx._adopt(self)

    def gmres(self, x0: DeviceArray, b: DeviceArray, z: DeviceArray, rtol: float, restart: int,
              max_iters: int, zero_guess: bool = True) -> GmresResult:
        r = GmresResult()
        check(self.ctx.lib.numa_hip_gmres(self.h, x0.ptr, b.ptr, z.ptr, rtol, restart, max_iters,
                                          1 if zero_guess else 0, C.byref(r)))
        return r

    def solve(self, params: KrylovParams, x: DeviceArray) -> KrylovResult:
        r = KrylovResult()
        check(self.ctx.lib.numa_hip_solve(self.h, C.byref(params), x.ptr, C.byref(r)))
        return r

    def solve_host(self, params: KrylovParams, x: np.ndarray) -> KrylovResult:
        assert x.dtype == np.float64 and x.flags.c_contiguous
        r = KrylovResult()
        check(self.ctx.lib.numa_hip_solve_host(self.h, C.byref(params),
                                               x.ctypes.data_as(_dp), C.byref(r)))
        return r

    def set_net(self, net: "ConvNet | None", fallback_ratio: float = 0.5) -> None:
        """Attach a learned preconditioner (PRECOND_LEARNED solvers)."""
        check(self.ctx.lib.numa_hip_solver_set_net(self.h, net.h if net is not None else None,
                                                   fallback_ratio))
        self._net = net

    def set_timing(self, on: bool = True) -> None:
        check(self.ctx.lib.numa_hip_solver_set_timing(self.h, 1 if on else 0))

    def timers(self, reset: bool = False) -> Timers:
        t = Timers()
        check(self.ctx.lib.numa_hip_solver_timers(self.h, C.byref(t), 1 if reset else 0))
        return t

    def set_graphs(self, on: bool = True) -> None:
        check(self.ctx.lib.numa_hip_solver_set_graphs(self.h, 1 if on else 0))

    def eigen_smallest(self, x: "DeviceArray", params: EigenParams | None = None,
                       vec: "DeviceArray | None" = None) -> EigenResult:
        """inverse_smallest_eigenpair (analysis/spectral.hpp:34-115) of J(x)."""
        r = EigenResult()
        p = params if params is not None else eigen_params()
        check(self.ctx.lib.numa_hip_eigen_smallest(self.h, x.ptr, C.byref(p), C.byref(r),
                                                   vec.ptr if vec is not None else None))
        return r

    def tf2018_curvature(self, x: "DeviceArray", d: "DeviceArray", out: "DeviceArray") -> float:
        """tensor_free_curvature_2018 (tf_chebyshev_2018.hpp:31-51) on device
        FP64 vectors; returns the step h."""
        h = C.c_double()
        check(self.ctx.lib.numa_hip_tf2018_curvature(self.h, x.ptr, d.ptr, out.ptr, C.byref(h)))
        return h.value

    def graphs_cached(self) -> int:
        c = _i64()
        check(self.ctx.lib.numa_hip_solver_graphs_cached(self.h, C.byref(c)))
        return int(c.value)

    def mg_fusion(self) -> list[tuple[int, int, int]]:
        """(NUMA_FUSE_* mode, plane chunk, rows per thread) per multigrid level."""
        L = _i64()
        check(self.ctx.lib.numa_hip_solver_mg_fusion(self.h, C.byref(L), None, None, None))
        k = max(1, L.value)
        m, ic, rw = (_i64 * k)(), (_i64 * k)(), (_i64 * k)()
        check(self.ctx.lib.numa_hip_solver_mg_fusion(self.h, C.byref(L), m, ic, rw))
        return [(int(m[i]), int(ic[i]), int(rw[i])) for i in range(L.value)]

    def set_sweep_events(self, on: bool = True) -> None:
        """Bracket every finest-level smoothing sweep that reads the iterate
        with HIP events (the solve then runs eagerly)."""
        check(self.ctx.lib.numa_hip_solver_sweep_events(self.h, 1 if on else 0))

    def sweep_time(self) -> tuple[float, int]:
        """(average ms, count) of the bracketed sweeps since the last call."""
        ms, cnt = C.c_double(), _i64()
        check(self.ctx.lib.numa_hip_solver_sweep_time(self.h, C.byref(ms), C.byref(cnt)))
        return ms.value, int(cnt.value)

    def inject_nan(self, what: int, call: int) -> None:
        """NaN into the call-th F evaluation (INJECT_F) or J v product
        (INJECT_JV) of the following solves; fires once; call = 0 disarms."""
        check(self.ctx.lib.numa_hip_solver_inject_nan(self.h, what, call))

    def learned_stats(self) -> dict:
        s = (_i64 * 3)()
        check(self.ctx.lib.numa_hip_solver_learned_stats(self.h, s))
        return {"net_applications": int(s[0]), "gmg_applications": int(s[1]),
                "fallbacks": int(s[2])}

    def primitive_cost(self, prim: int, reps: int = 20) -> tuple[float, float]:
        """(ms per call, algorithmic bytes per call) of one primitive."""
        ms, b = C.c_double(), C.c_double()
        check(self.ctx.lib.numa_hip_primitive_cost(self.h, prim, reps, C.byref(ms), C.byref(b)))
        return ms.value, b.value

    def close(self):
        if self.h and self.ctx.h:
            check(self.ctx.lib.numa_hip_solver_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def newton_krylov(ctx, n, lam, x0: np.ndarray, params=None, cfg=None, restart=None):
    """numa::op::newton_krylov (krylov_newton.hpp:69-114) on the device."""
    cfg = cfg if cfg is not None else solver_config()
    cfg.method = METHOD_NEWTON
    params = params if params is not None else krylov_params()
    s = BratuSolver(ctx, n, lam, cfg, restart or max(1, params.gmres_restart))
    x = np.array(x0, np.float64)
    r = s.solve_host(params, x)
    s.close()
    return x, r


def chebyshev_inexact(ctx, n, lam, x0: np.ndarray, params=None, cfg=None, restart=None):
    """numa::op::chebyshev_inexact (krylov_newton.hpp:116-189) on the device."""
    cfg = cfg if cfg is not None else solver_config()
    cfg.method = METHOD_CHEBYSHEV
    params = params if params is not None else krylov_params()
    s = BratuSolver(ctx, n, lam, cfg, restart or max(1, params.gmres_restart))
    x = np.array(x0, np.float64)
    r = s.solve_host(params, x)
    s.close()
    return x, r


def continuation(ctx, n: int, cfg=None, cp: ContinuationParams | None = None, nk=None,
                 max_points: int = 512, x_out: "DeviceArray | None" = None):
    """apps/calibrate_b6.cpp:173-298 on the device: returns (points, status,
    fold proxy); status 0 ok, 3 corrector failed below min_step, 4 spectral
    estimate failed.  cfg defaults to calibrate_b6's solver (analytic J v,
    PoissonMG(n))."""
    if cfg is None:
        cfg = solver_config(precond=PRECOND_GMG, jv_mode=DERIV_ANALYTIC, mg=mg_reference())
    cp = cp if cp is not None else continuation_params()
    nk = nk if nk is not None else corrector_params()
    pts = (ContinuationPoint * max_points)()
    count, status, proxy = _i64(), _i64(), C.c_double()
    check(ctx.lib.numa_hip_continuation(ctx.h, int(n), C.byref(cfg), C.byref(cp), C.byref(nk),
                                        pts, max_points, C.byref(count), C.byref(status),
                                        C.byref(proxy),
                                        x_out.ptr if x_out is not None else None))
    return list(pts[:min(count.value, max_points)]), int(status.value), proxy.value


# ---- start vectors (C1 campaign files) ---------------------------------------
def write_start(path: str, x: np.ndarray) -> None:
    """One value per line, 17 significant digits (materialize_coarse_start.cpp:18-43)."""
    x = np.ascontiguousarray(x, np.float64)
    check(load().numa_hip_start_write(os.fsencode(path), x.ctypes.data, x.size))


def read_start(path: str, expected: int) -> np.ndarray:
    """load_vector (run_campaign.cpp:287-302): finite values, exactly `expected`."""
    x = np.empty(int(expected))
    check(load().numa_hip_start_read(os.fsencode(path), int(expected), x.ctypes.data))
    return x


def sha256_bytes(data: bytes) -> str:
    out = C.create_string_buffer(65)
    buf = C.create_string_buffer(bytes(data), len(data))
    check(load().numa_hip_sha256(buf, len(data), out))
    return out.value.decode()


def sha256_file(path: str) -> str:
    """The *_sha256 columns of the frozen start manifests (experiments/c1/starts/frozen)."""
    out = C.create_string_buffer(65)
    check(load().numa_hip_sha256_file(os.fsencode(path), out))
    return out.value.decode()


def perturb(ref: np.ndarray, rho: float, seed: int) -> np.ndarray:
    """perturb (run_campaign.cpp:304-327): ref + rho ||ref|| d / ||d||, d ~ N(0, 1)
    from std::mt19937_64(seed * 7919 + 13)."""
    ref = np.ascontiguousarray(ref, np.float64)
    out = np.empty_like(ref)
    check(load().numa_hip_perturb_host(ref.ctypes.data, ref.size, float(rho), int(seed),
                                       out.ctypes.data))
    return out


def prolong3d(ctx: "Context", coarse: "DeviceArray", nc: int, fine: "DeviceArray") -> None:
    """B1 coarse start (materialize_coarse_start.cpp:60-167) on the device:
    fine (size (2 nc + 1)^3) = trilinear prolongation of coarse (nc^3)."""
    check(ctx.lib.numa_hip_prolong3d(ctx.h, coarse.ptr, int(nc), fine.ptr))


def tf_chebyshev_2018(ctx, n, lam, x0: np.ndarray, params=None, cfg=None, restart=None):
    """numa::op::tf_chebyshev_2018 (tf_chebyshev_2018.hpp:53-132) on the device."""
    cfg = cfg if cfg is not None else solver_config()
    cfg.method = METHOD_TF2018
    params = params if params is not None else krylov_params()
    s = BratuSolver(ctx, n, lam, cfg, restart or max(1, params.gmres_restart))
    x = np.array(x0, np.float64)
    r = s.solve_host(params, x)
    s.close()
    return x, r


def chandra_quadrature(n: int, rule: int = 0):
    """Nodes and weights on [0, 1]: Gauss-Legendre (rule 0) or midpoint (1)."""
    load()
    a, w = np.empty(n), np.empty(n)
    check(_lib.numa_hip_chandra_quadrature(n, rule, a.ctypes.data_as(_dp), w.ctypes.data_as(_dp)))
    return a, w


class ChandrasekharBatch:
    """problems/chandrasekhar_op.hpp:31-136 for a batch of parameters c on the
    device (rule 0 Gauss-Legendre / 1 midpoint; form 0 reference / 1 Kelley),
    solved by lockstep newton_krylov (krylov_newton.hpp:69-114) or, with
    method = METHOD_CHEBYSHEV, chebyshev_inexact (:116-189).
    Vectors are instance-major numpy arrays of shape (batch, n)."""

    def __init__(self, ctx: Context, n: int, cvals, rule: int = 0, form: int = 0,
                 precision: int = PRECISION_FP64, restart: int = 50,
                 method: int = METHOD_NEWTON):
        self.ctx, self.n = ctx, int(n)
        self.c = np.ascontiguousarray(cvals, np.float64).ravel()
        self.batch = self.c.size
        h = _vp()
        check(ctx.lib.numa_hip_chandra_create(ctx.h, self.n, self.batch,
                                              self.c.ctypes.data_as(_dp), rule, form, precision,
                                              int(restart), C.byref(h)))
        self.h = h
        ctx._adopt(self)
        if method != METHOD_NEWTON:
            check(ctx.lib.numa_hip_chandra_set_method(self.h, method))

    def _arr(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(self.batch, self.n)
        return self.ctx.array(x.ravel())

    def eval(self, x):
        dx = self._arr(x)
        dF = self.ctx.empty(self.batch * self.n)
        norms = np.empty(self.batch)
        check(self.ctx.lib.numa_hip_chandra_eval(self.h, dx.ptr, dF.ptr, norms.ctypes.data_as(_dp)))
        return dF.numpy().reshape(self.batch, self.n), norms

    def jvp(self, x, v):
        dx, dv = self._arr(x), self._arr(v)
        do = self.ctx.empty(self.batch * self.n)
        check(self.ctx.lib.numa_hip_chandra_jvp(self.h, dx.ptr, dv.ptr, do.ptr))
        return do.numpy().reshape(self.batch, self.n)

    def d2(self, x, v):
        dx, dv = self._arr(x), self._arr(v)
        do = self.ctx.empty(self.batch * self.n)
        check(self.ctx.lib.numa_hip_chandra_d2(self.h, dx.ptr, dv.ptr, do.ptr))
        return do.numpy().reshape(self.batch, self.n)

    def solve(self, params: KrylovParams, x0=None):
        x0 = np.ones((self.batch, self.n)) if x0 is None else x0
        dx = self._arr(x0)
        res = (KrylovResult * self.batch)()
        check(self.ctx.lib.numa_hip_chandra_solve(self.h, C.byref(params), dx.ptr, res))
        return dx.numpy().reshape(self.batch, self.n), list(res)

    def solve_device(self, params: KrylovParams, dx: "DeviceArray"):
        res = (KrylovResult * self.batch)()
        check(self.ctx.lib.numa_hip_chandra_solve(self.h, C.byref(params), dx.ptr, res))
        return list(res)

    def gemm_cost(self, precision: int, reps: int = 10):
        ms, fl = C.c_double(), C.c_double()
        check(self.ctx.lib.numa_hip_chandra_gemm_cost(self.h, precision, reps, C.byref(ms),
                                                      C.byref(fl)))
        return ms.value, fl.value

    def close(self):
        if self.h and self.ctx.h:
            check(self.ctx.lib.numa_hip_chandra_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def net_train_config(pairs: int = 2000, batch: int = 4, sizes=(32, 48, 64), lambda_lo: float = 1.0,
                     lambda_hi: float = 6.7, noise: float = 0.1, lr: float = 1e-3, seed: int = 7,
                     gmres_restart: int = 30, mg: MGConfig | None = None,
                     solve: KrylovParams | None = None, loss: int = LOSS_RESIDUAL,
                     lr_end: float = 0.0) -> NetTrainConfig:
    sz = (_i64 * 4)(*([int(v) for v in sizes] + [0] * (4 - len(sizes))))
    return NetTrainConfig(pairs=pairs, batch=batch, sizes=sz, lambda_lo=lambda_lo,
                          lambda_hi=lambda_hi, noise=noise, lr=lr, seed=seed,
                          gmres_restart=gmres_restart,
                          mg=mg if mg is not None else mg_cheb(levels=5),
                          solve=solve if solve is not None else krylov_params(gmres_restart=30),
                          loss=loss, lr_end=lr_end)


class ConvNet:
    """The learned preconditioner's network (numa_hip_net): 6 conv layers
    (1-24-24-24-24-24-1, 3x3x3, GELU), bf16 MFMA on the 24->24 layers."""

    def __init__(self, ctx: Context, seed: int = 1):
        self.ctx = ctx
        h = _vp()
        check(ctx.lib.numa_hip_net_create(ctx.h, int(seed), C.byref(h)))
        self.h = h
        cnt = _i64()
        check(ctx.lib.numa_hip_net_param_count(C.byref(cnt)))
        self.n_params = int(cnt.value)
        ctx._adopt(self)

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(self.ctx.lib.numa_hip_net_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.n_params
        check(self.ctx.lib.numa_hip_net_set_params(self.h, w.ctypes.data))

    def forward(self, n: int, x: np.ndarray) -> np.ndarray:
        """Raw network on a batch of grids, x of shape (B, n, n, n)."""
        x = np.ascontiguousarray(x, np.float32)
        B = x.size // n ** 3
        dx = self.ctx.array(x.ravel())
        do = self.ctx.empty(x.size, np.float32)
        check(self.ctx.lib.numa_hip_net_forward(self.h, n, B, dx.ptr, do.ptr))
        return do.numpy().reshape(x.shape)

    def loss_grad(self, n: int, x: np.ndarray, y: np.ndarray,
                  loss: int = LOSS_CORRECTION) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        B = x.size // n ** 3
        dx, dy = self.ctx.array(x.ravel()), self.ctx.array(y.ravel())
        g = np.empty(self.n_params, np.float32)
        lv = C.c_double()
        check(self.ctx.lib.numa_hip_net_loss_grad(self.h, n, B, dx.ptr, dy.ptr, loss, C.byref(lv),
                                                  g.ctypes.data))
        return lv.value, g

    def train_step(self, n: int, r: "DeviceArray", z: "DeviceArray", batch: int,
                   lr: float, loss: int = LOSS_RESIDUAL) -> float:
        lv = C.c_double()
        check(self.ctx.lib.numa_hip_net_train_step(self.h, n, batch, r.ptr, z.ptr, loss, lr,
                                                   C.byref(lv)))
        return lv.value

    def apply(self, n: int, r: "DeviceArray", z: "DeviceArray") -> None:
        check(self.ctx.lib.numa_hip_net_apply(self.h, n, r.ptr, z.ptr))

    def train_solver(self, cfg: NetTrainConfig, hist_len: int = 0):
        hist = np.zeros(max(1, hist_len))
        st = NetTrainStats()
        check(self.ctx.lib.numa_hip_net_train_solver(self.h, C.byref(cfg),
                                                     hist.ctypes.data_as(_dp), hist_len,
                                                     C.byref(st)))
        return st, hist[:min(hist_len, st.steps)]

    def cost(self, n: int, what: int = 0, batch: int = 1, reps: int = 5) -> float:
        ms = C.c_double()
        check(self.ctx.lib.numa_hip_net_cost(self.h, n, what, batch, reps, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h and self.ctx.h:
            check(self.ctx.lib.numa_hip_net_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uniform_host(seed: int, count: int, scale: float = 1.0) -> np.ndarray:
    """scale * U[0,1) from std::mt19937_64(seed), U = (u64 >> 11) 2^-53 (host)."""
    L = load()
    out = np.empty(int(count))
    check(L.numa_hip_uniform_host(int(seed), int(count), float(scale), out.ctypes.data_as(_dp)))
    return out


def sweep_instances(seed: int, count: int, sizes=(32, 48, 64, 96), lambda_lo: float = 1.0,
                    lambda_hi: float = 6.7) -> tuple[np.ndarray, np.ndarray]:
    """Seeded (n, lambda) of a sweep (std::mt19937_64, see include/numa_hip.h)."""
    L = load()
    sz = np.ascontiguousarray(sizes, np.int64)
    n = np.empty(int(count), np.int64)
    lam = np.empty(int(count))
    check(L.numa_hip_sweep_instances(int(seed), int(count),
                                     sz.ctypes.data_as(C.POINTER(_i64)), sz.size, lambda_lo,
                                     lambda_hi, n.ctypes.data_as(C.POINTER(_i64)),
                                     lam.ctypes.data_as(_dp)))
    return n, lam


class Sweep:
    """Batched sweep over Bratu3D instances (numa_hip_sweep): `workers`
    streams with per-size solvers, driven by host threads."""

    def __init__(self, ctx: Context, cfg: SolverConfig, workers: int = 8, restart: int = 30):
        self.ctx, self.cfg = ctx, cfg
        h = _vp()
        check(ctx.lib.numa_hip_sweep_create(ctx.h, int(workers), C.byref(cfg), int(restart),
                                            C.byref(h)))
        self.h = h
        ctx._adopt(self)

    def run(self, ns, lambdas, xs: list, params: KrylovParams, graphs: bool = True):
        """Solves instance i from xs[i] (DeviceArray, overwritten); returns
        (list of KrylovResult, wall seconds)."""
        count = len(xs)
        n = np.ascontiguousarray(ns, np.int64)
        lam = np.ascontiguousarray(lambdas, np.float64)
        ptrs = (_vp * max(1, count))(*[x.ptr for x in xs])
        res = (KrylovResult * max(1, count))()
        sec = C.c_double()
        check(self.ctx.lib.numa_hip_sweep_run(self.h, count, n.ctypes.data_as(C.POINTER(_i64)),
                                              lam.ctypes.data_as(_dp), ptrs, C.byref(params),
                                              1 if graphs else 0, res, C.byref(sec)))
        return list(res)[:count], sec.value

    def close(self):
        if self.h and self.ctx.h:
            check(self.ctx.lib.numa_hip_sweep_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched solver (numa_hip_batch): instances (n_i, lambda_i) solved
    together, one launch per operation over the instance table and one
    (graph) launch per lockstep Arnoldi step."""

    def __init__(self, ctx: Context, ns, lambdas, cfg: SolverConfig, restart: int = 30):
        self.ctx, self.cfg = ctx, cfg
        self.n = np.ascontiguousarray(ns, np.int64)
        lam = np.ascontiguousarray(lambdas, np.float64)
        if self.n.shape != lam.shape:
            raise ValueError("ns and lambdas differ in length")
        h = _vp()
        check(ctx.lib.numa_hip_batch_create(ctx.h, len(self.n),
                                            self.n.ctypes.data_as(C.POINTER(_i64)),
                                            lam.ctypes.data_as(_dp), C.byref(cfg), int(restart),
                                            C.byref(h)))
        self.h = h
        ctx._adopt(self)

    def set_graphs(self, on: bool = True):
        check(self.ctx.lib.numa_hip_batch_set_graphs(self.h, 1 if on else 0))

    def set_smoothing(self, forms):
        """Pins the smoothing form of levels 0.. (0 per step, 1..3 box shape)."""
        f = np.ascontiguousarray(forms, np.int64)
        check(self.ctx.lib.numa_hip_batch_set_smoothing(self.h, len(f),
                                                        f.ctypes.data_as(C.POINTER(_i64))))

    def solve(self, xs: list, params: KrylovParams):
        """Solves instance i from xs[i] (DeviceArray of n_i^3, overwritten);
        returns (list of KrylovResult, wall seconds)."""
        count = len(self.n)
        if len(xs) != count:
            raise ValueError(f"expected {count} iterates, got {len(xs)}")
        for x, n in zip(xs, self.n):
            if x.n != int(n) ** 3:
                raise ValueError("iterate length does not match its grid")
        ptrs = (_vp * count)(*[x.ptr for x in xs])
        res = (KrylovResult * count)()
        sec = C.c_double()
        check(self.ctx.lib.numa_hip_batch_solve(self.h, C.byref(params), ptrs, res,
                                                C.byref(sec)))
        return list(res), sec.value

    def stats(self, reset: bool = False) -> dict:
        st = BatchStats()
        check(self.ctx.lib.numa_hip_batch_stats(self.h, 1 if reset else 0, C.byref(st)))
        return st.as_dict()

    def close(self):
        if self.h and self.ctx.h:
            check(self.ctx.lib.numa_hip_batch_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
