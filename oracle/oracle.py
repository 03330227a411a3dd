"""ctypes bindings of the CPU oracle (liboracle.so) and of the reference's own
header library compiled into oracle/_ref/libnuma_ref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference results are compared with or the
timed CPU baseline.  It is never
a product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libnuma_ref.so")

_dp = C.POINTER(C.c_double)
_i64 = C.c_int64


# ---- POD structs (include/numa_hip.h) -------------------------------------
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


class GmresResult(C.Structure):
    _fields_ = [("iters", _i64), ("matvecs", _i64), ("precs", _i64), ("rel_res", C.c_double),
                ("converged", _i64), ("breakdown", _i64)]


class EigenResult(C.Structure):
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
    _fields_ = [("lambda_", C.c_double), ("spectral_gap", C.c_double),
                ("eigen_residual", C.c_double), ("continuation_iters", _i64), ("n_F", _i64),
                ("n_JVP", _i64), ("n_prec", _i64)]


def eigen_params(eigen_rtol=1e-7, linear_rtol=1e-10, max_outer=30, restart=100,
                 max_linear=2000) -> EigenParams:
    """analysis/spectral.hpp:34-42 defaults."""
    return EigenParams(eigen_rtol, linear_rtol, max_outer, restart, max_linear)


def continuation_params(**kw) -> ContinuationParams:
    """apps/calibrate_b6.cpp:30-40 defaults."""
    d = dict(lambda_start=0.0, lambda_stop=9.8, step=0.25, min_step=1e-4, spectral_stop=5e-2,
             eigen_rtol=1e-7, linear_rtol=1e-10)
    d.update(kw)
    return ContinuationParams(**d)


def corrector_params() -> "KrylovParams":
    """The continuation corrector's Newton-Krylov parameters (calibrate_b6.cpp:181-187)."""
    return krylov_params(rtol_nl=1e-11, max_nl=100, eta_init=1e-8, eta_max=1e-6,
                         gmres_restart=100, gmres_max_iters=2000)


class MGConfig(C.Structure):
    _fields_ = [("smoother", _i64), ("pre", _i64), ("post", _i64), ("degree", _i64),
                ("max_levels", _i64), ("min_coarse_n", _i64), ("coarse_iters", _i64),
                ("even_coarsening", _i64), ("omega", C.c_double), ("cheb_ratio", C.c_double),
                ("fuse", _i64), ("ichunk", _i64), ("rows", _i64)]


class SolverConfig(C.Structure):
    _fields_ = [("method", _i64), ("precond", _i64), ("precision", _i64), ("flexible", _i64),
                ("jv_mode", _i64), ("d2_mode", _i64), ("mg", MGConfig)]


METHOD_NEWTON, METHOD_CHEBYSHEV, METHOD_TF2018 = 0, 1, 2
PRECOND_NONE, PRECOND_JACOBI, PRECOND_GMG = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
DERIV_FD, DERIV_ANALYTIC = 0, 1
FORCING_EW, FORCING_CONSTANT = 0, 1
SMOOTH_RBGS, SMOOTH_WJACOBI, SMOOTH_CHEBYSHEV = 0, 1, 2


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
    """PoissonMG(n) of linalg/poisson_mg.hpp:26-64: RB-GS V(2,2), odd-n coarsening."""
    return MGConfig(smoother=SMOOTH_RBGS, pre=2, post=2, degree=3, max_levels=0, min_coarse_n=1,
                    coarse_iters=30, even_coarsening=0, omega=6.0 / 7.0, cheb_ratio=4.0)


def mg_config(**kw) -> MGConfig:
    m = mg_reference()
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def solver_config(method=METHOD_NEWTON, precond=PRECOND_GMG, precision=PRECISION_FP64,
                  flexible=0, jv_mode=DERIV_FD, d2_mode=DERIV_FD, mg=None) -> SolverConfig:
    return SolverConfig(method=method, precond=precond, precision=precision, flexible=flexible,
                        jv_mode=jv_mode, d2_mode=d2_mode, mg=mg if mg is not None else mg_reference())


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        self.lib = L
        for name in ("orc_bratu_eval", "orc_bratu_jacobi_diag"):
            getattr(L, name).argtypes = [_i64, C.c_double, _dp, _dp]
        for name in ("orc_bratu_jvp", "orc_bratu_d2", "orc_bratu_jvp_fd", "orc_bratu_d2_fd"):
            getattr(L, name).argtypes = [_i64, C.c_double, _dp, _dp, _dp]
        L.orc_poisson_apply.argtypes = [_i64, _dp, _dp]
        L.orc_dot.argtypes = [_i64, _dp, _dp]
        L.orc_dot.restype = C.c_double
        L.orc_mg_levels.argtypes = [_i64, C.POINTER(MGConfig), C.POINTER(_i64)]
        L.orc_mg_levels.restype = _i64
        L.orc_mg_apply.argtypes = [_i64, C.POINTER(MGConfig), _dp, _dp]
        L.orc_bratu_gmres.argtypes = [_i64, C.c_double, C.POINTER(SolverConfig), _dp, _dp, _dp,
                                      C.c_double, _i64, _i64, _i64, C.POINTER(GmresResult)]
        L.orc_poisson_gmres.argtypes = [_i64, C.POINTER(MGConfig), _i64, _dp, _dp, C.c_double,
                                        _i64, _i64, C.POINTER(GmresResult)]
        L.orc_bratu_solve.argtypes = [_i64, C.c_double, C.POINTER(SolverConfig),
                                      C.POINTER(KrylovParams), _dp, C.POINTER(KrylovResult)]
        L.orc_tf2018_curvature.argtypes = [_i64, C.c_double, _i64, _dp, _dp, _dp, _dp]
        L.orc_bratu_eigen_smallest.argtypes = [_i64, C.c_double, C.POINTER(SolverConfig), _dp,
                                               C.POINTER(EigenParams), C.POINTER(EigenResult),
                                               _dp]
        L.orc_bratu_continuation.argtypes = [_i64, C.POINTER(SolverConfig),
                                             C.POINTER(ContinuationParams),
                                             C.POINTER(KrylovParams),
                                             C.POINTER(ContinuationPoint), _i64,
                                             C.POINTER(_i64), C.POINTER(_i64),
                                             C.POINTER(C.c_double), _dp]
        L.orc_gauss_legendre_unit.argtypes = [_i64, _dp, _dp]
        L.orc_chandra_eval.argtypes = [_i64, C.c_double, _i64, _i64, _dp, _dp]
        L.orc_chandra_jvp.argtypes = [_i64, C.c_double, _i64, _i64, _dp, _dp, _dp]
        L.orc_chandra_d2.argtypes = [_i64, C.c_double, _i64, _i64, _dp, _dp, _dp]
        L.orc_chandra_solve.argtypes = [_i64, C.c_double, _i64, _i64, C.POINTER(SolverConfig),
                                        C.POINTER(KrylovParams), _dp, C.POINTER(KrylovResult)]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_inject.argtypes = [_i64, _i64]
        L.orc_set_inject.restype = None

    def set_inject(self, what: int, at: int) -> None:
        """NaN into the at-th F evaluation (what 0) or J v product (what 1) of
        the following Bratu solves; fires once; at = 0 disarms."""
        self.lib.orc_set_inject(what, at)

    def threads(self) -> int:
        return int(self.lib.orc_num_threads())

    # Bratu3D ---------------------------------------------------------------
    def bratu_eval(self, n, lam, u):
        u = np.ascontiguousarray(u, np.float64)
        out = np.empty(n ** 3)
        self.lib.orc_bratu_eval(n, lam, _ptr(u), _ptr(out))
        return out

    def _uv(self, name, n, lam, u, v):
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        out = np.empty(n ** 3)
        getattr(self.lib, name)(n, lam, _ptr(u), _ptr(v), _ptr(out))
        return out

    def bratu_jvp(self, n, lam, u, v):
        return self._uv("orc_bratu_jvp", n, lam, u, v)

    def bratu_d2(self, n, lam, u, v):
        return self._uv("orc_bratu_d2", n, lam, u, v)

    def bratu_jvp_fd(self, n, lam, x, v):
        return self._uv("orc_bratu_jvp_fd", n, lam, x, v)

    def bratu_d2_fd(self, n, lam, x, v):
        return self._uv("orc_bratu_d2_fd", n, lam, x, v)

    def bratu_jacobi_diag(self, n, lam, u):
        u = np.ascontiguousarray(u, np.float64)
        out = np.empty(n ** 3)
        self.lib.orc_bratu_jacobi_diag(n, lam, _ptr(u), _ptr(out))
        return out

    def poisson_apply(self, n, v):
        v = np.ascontiguousarray(v, np.float64)
        out = np.empty(n ** 3)
        self.lib.orc_poisson_apply(n, _ptr(v), _ptr(out))
        return out

    def dot(self, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        return float(self.lib.orc_dot(a.size, _ptr(a), _ptr(b)))

    # multigrid ---------------------------------------------------------------
    def mg_levels(self, n, cfg=None):
        sizes = (_i64 * 64)()
        L = self.lib.orc_mg_levels(n, C.byref(cfg) if cfg else None, sizes)
        return [int(sizes[i]) for i in range(L)]

    def mg_apply(self, n, r, cfg=None):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(n ** 3)
        rc = self.lib.orc_mg_apply(n, C.byref(cfg) if cfg else None, _ptr(r), _ptr(z))
        assert rc == 0
        return z

    # Krylov ------------------------------------------------------------------
    def bratu_gmres(self, n, lam, cfg, x0, b, rtol, restart, max_iters, x=None, zero_guess=1):
        x0 = np.ascontiguousarray(x0, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n ** 3) if x is None else np.array(x, np.float64)
        res = GmresResult()
        rc = self.lib.orc_bratu_gmres(n, lam, C.byref(cfg), _ptr(x0), _ptr(b), _ptr(x), rtol,
                                      restart, max_iters, zero_guess, C.byref(res))
        assert rc == 0
        return x, res

    def poisson_gmres(self, n, b, rtol, restart, max_iters, mg=None, use_mg=1):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n ** 3)
        res = GmresResult()
        self.lib.orc_poisson_gmres(n, C.byref(mg) if mg else None, use_mg, _ptr(b), _ptr(x),
                                   rtol, restart, max_iters, C.byref(res))
        return x, res

    def bratu_solve(self, n, lam, cfg, params, x0=None):
        x = np.zeros(n ** 3) if x0 is None else np.array(x0, np.float64)
        res = KrylovResult()
        rc = self.lib.orc_bratu_solve(n, lam, C.byref(cfg), C.byref(params), _ptr(x), C.byref(res))
        assert rc == 0
        return x, res

    def bratu_eigen_smallest(self, n, lam, cfg, x, params=None):
        """inverse_smallest_eigenpair (analysis/spectral.hpp:34-115) of J(x)."""
        x = np.ascontiguousarray(x, np.float64)
        p = params if params is not None else eigen_params()
        res, vec = EigenResult(), np.zeros(n ** 3)
        assert self.lib.orc_bratu_eigen_smallest(n, lam, C.byref(cfg), _ptr(x), C.byref(p),
                                                 C.byref(res), _ptr(vec)) == 0
        return res, vec

    def bratu_continuation(self, n, cfg, cp, nk=None, max_points=512):
        """apps/calibrate_b6.cpp:173-298; returns (points, status, fold proxy, last x)."""
        nk = nk if nk is not None else corrector_params()
        pts = (ContinuationPoint * max_points)()
        count, status, proxy = _i64(), _i64(), C.c_double()
        x = np.zeros(n ** 3)
        assert self.lib.orc_bratu_continuation(n, C.byref(cfg), C.byref(cp), C.byref(nk), pts,
                                               max_points, C.byref(count), C.byref(status),
                                               C.byref(proxy), _ptr(x)) == 0
        return list(pts[:min(count.value, max_points)]), int(status.value), proxy.value, x

    def tf2018_curvature(self, n, lam, x, d, fd=0):
        """tensor_free_curvature_2018 (tf_chebyshev_2018.hpp:31-51) with the
        analytic (fd = 0) or finite-difference (fd = 1) J v; returns (C d, h)."""
        x = np.ascontiguousarray(x, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        out = np.empty(n ** 3)
        h = C.c_double()
        mode = DERIV_FD if fd else DERIV_ANALYTIC
        assert self.lib.orc_tf2018_curvature(n, lam, mode, _ptr(x), _ptr(d), _ptr(out),
                                             C.byref(h)) == 0
        return out, h.value

    # Chandrasekhar -------------------------------------------------------------
    def gauss_legendre_unit(self, n):
        a, b = np.empty(n), np.empty(n)
        self.lib.orc_gauss_legendre_unit(n, _ptr(a), _ptr(b))
        return a, b

    def chandra_eval(self, n, c, x, rule=0, form=0):
        x = np.ascontiguousarray(x, np.float64)
        o = np.empty(n)
        self.lib.orc_chandra_eval(n, c, rule, form, _ptr(x), _ptr(o))
        return o

    def chandra_jvp(self, n, c, x, v, rule=0, form=0):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        o = np.empty(n)
        self.lib.orc_chandra_jvp(n, c, rule, form, _ptr(x), _ptr(v), _ptr(o))
        return o

    def chandra_d2(self, n, c, x, v, rule=0, form=0):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        o = np.empty(n)
        self.lib.orc_chandra_d2(n, c, rule, form, _ptr(x), _ptr(v), _ptr(o))
        return o

    def chandra_solve(self, n, c, cfg, params, x0=None, rule=0, form=0):
        x = np.ones(n) if x0 is None else np.array(x0, np.float64)
        res = KrylovResult()
        self.lib.orc_chandra_solve(n, c, rule, form, C.byref(cfg), C.byref(params), _ptr(x),
                                   C.byref(res))
        return x, res


def params_vector(p: KrylovParams) -> np.ndarray:
    """KrylovParams in the 13-double layout ref_driver.cpp reads."""
    return np.array([p.rtol_nl, p.max_nl, p.eta_init, p.eta_max, p.gamma, p.alpha,
                     p.eta2_factor, p.eta2_max, p.gmres_restart, p.gmres_max_iters, p.armijo,
                     p.min_lambda, p.recycle], np.float64)


class Reference:
    """The reference's own headers compiled into oracle/_ref (this container only)."""

    def __init__(self, path: str = REF_SO):
        L = C.CDLL(path)
        self.lib = L
        for name in ("ref_bratu_eval", "ref_bratu_jacobi_diag"):
            getattr(L, name).argtypes = [_i64, C.c_double, _dp, _dp]
        for name in ("ref_bratu_jvp", "ref_bratu_d2", "ref_bratu_jvp_fd", "ref_bratu_d2_fd"):
            getattr(L, name).argtypes = [_i64, C.c_double, _dp, _dp, _dp]
        L.ref_mg_levels.argtypes = [_i64]
        L.ref_mg_levels.restype = _i64
        L.ref_mg_apply.argtypes = [_i64, _dp, _dp]
        L.ref_bratu_gmres.argtypes = [_i64, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp,
                                      C.c_double, _i64, _i64, _dp]
        L.ref_bratu_solve.argtypes = [_i64, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.ref_bratu_solve_inject.argtypes = [_i64, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp,
                                             C.c_int, _i64]
        L.ref_tf2018_curvature.argtypes = [_i64, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
        L.ref_bratu_eigen_smallest.argtypes = [_i64, C.c_double, _dp, C.c_double, C.c_double,
                                               _i64, _i64, _i64, _dp, _dp]
        L.ref_gauss_legendre_unit.argtypes = [_i64, _dp, _dp]
        L.ref_chandra_eval.argtypes = [_i64, C.c_double, _dp, _dp]
        L.ref_chandra_jvp.argtypes = [_i64, C.c_double, _dp, _dp, _dp]
        L.ref_chandra_d2.argtypes = [_i64, C.c_double, _dp, _dp, _dp]
        L.ref_chandra_solve.argtypes = [_i64, C.c_double, C.c_int, _dp, _dp, _dp]

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    def bratu(self, name, n, lam, *vecs):
        vecs = [np.ascontiguousarray(v, np.float64) for v in vecs]
        out = np.empty(n ** 3)
        getattr(self.lib, "ref_bratu_" + name)(n, lam, *[_ptr(v) for v in vecs], _ptr(out))
        return out

    def mg_levels(self, n):
        return int(self.lib.ref_mg_levels(n))

    def mg_apply(self, n, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(n ** 3)
        self.lib.ref_mg_apply(n, _ptr(r), _ptr(z))
        return z

    def bratu_gmres(self, n, lam, fd, precond, x0, b, rtol, restart, max_iters):
        x0 = np.ascontiguousarray(x0, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n ** 3)
        res = np.zeros(6)
        self.lib.ref_bratu_gmres(n, lam, fd, precond, _ptr(x0), _ptr(b), _ptr(x), rtol, restart,
                                 max_iters, _ptr(res))
        return x, dict(iters=int(res[0]), matvecs=int(res[1]), precs=int(res[2]),
                       rel_res=float(res[3]), converged=bool(res[4]), breakdown=bool(res[5]))

    def bratu_solve(self, n, lam, method, precond, fd, params: KrylovParams, x0=None):
        x = np.zeros(n ** 3) if x0 is None else np.array(x0, np.float64)
        pv = params_vector(params)
        res = np.zeros(10)
        self.lib.ref_bratu_solve(n, lam, method, precond, fd, _ptr(pv), _ptr(x), _ptr(res))
        return x, _result_dict(res)

    def bratu_solve_inject(self, n, lam, method, precond, params: KrylovParams, what, at):
        """Analytic-derivative solve from u0 = 0 with NaN in element 0 of the
        at-th F evaluation (what 0) or J v product (what 1)."""
        x = np.zeros(n ** 3)
        pv = params_vector(params)
        res = np.zeros(10)
        self.lib.ref_bratu_solve_inject(n, lam, method, precond, _ptr(pv), _ptr(x), _ptr(res),
                                        what, at)
        return x, _result_dict(res)

    def bratu_eigen_smallest(self, n, lam, x, eigen_rtol=1e-7, linear_rtol=1e-10, max_outer=30,
                             restart=100, max_linear=2000):
        x = np.ascontiguousarray(x, np.float64)
        res, vec = np.zeros(8), np.zeros(n ** 3)
        self.lib.ref_bratu_eigen_smallest(n, lam, _ptr(x), eigen_rtol, linear_rtol, max_outer,
                                          restart, max_linear, _ptr(res), _ptr(vec))
        return dict(converged=bool(res[0]), eigenvalue=float(res[1]),
                    residual_rel=float(res[2]), outer_iterations=int(res[3]),
                    linear_iterations=int(res[4]), linear_matvecs=int(res[5]),
                    preconditioner_applications=int(res[6]),
                    linear_failures=int(res[7])), vec

    def tf2018_curvature(self, n, lam, x, d, fd=0):
        x = np.ascontiguousarray(x, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        out = np.empty(n ** 3)
        h = C.c_double()
        self.lib.ref_tf2018_curvature(n, lam, fd, _ptr(x), _ptr(d), _ptr(out), C.byref(h))
        return out, h.value

    def gauss_legendre_unit(self, n):
        a, b = np.empty(n), np.empty(n)
        self.lib.ref_gauss_legendre_unit(n, _ptr(a), _ptr(b))
        return a, b

    def chandra(self, name, n, c, *vecs):
        vecs = [np.ascontiguousarray(v, np.float64) for v in vecs]
        out = np.empty(n)
        getattr(self.lib, "ref_chandra_" + name)(n, c, *[_ptr(v) for v in vecs], _ptr(out))
        return out

    def chandra_solve(self, n, c, method, params: KrylovParams, x0=None):
        x = np.ones(n) if x0 is None else np.array(x0, np.float64)
        pv = params_vector(params)
        res = np.zeros(10)
        self.lib.ref_chandra_solve(n, c, method, _ptr(pv), _ptr(x), _ptr(res))
        return x, _result_dict(res)


def _result_dict(r):
    return dict(converged=bool(r[0]), iters=int(r[1]), n_F=int(r[2]), n_JVP=int(r[3]),
                n_D2=int(r[4]), n_lin_iters=int(r[5]), n_lin_failures=int(r[6]),
                n_prec=int(r[7]), res0=float(r[8]), res=float(r[9]))


def result_dict(r: KrylovResult) -> dict:
    return dict(converged=bool(r.converged), iters=int(r.iters), n_F=int(r.n_F),
                n_JVP=int(r.n_JVP), n_D2=int(r.n_D2), n_lin_iters=int(r.n_lin_iters),
                n_lin_failures=int(r.n_lin_failures), n_prec=int(r.n_prec), res0=float(r.res0),
                res=float(r.res))


def sin_point(n: int, scale: float) -> np.ndarray:
    """x_i = scale sin(0.37 i + 0.5), the reference tests' deterministic input
    (tests/test_operators.cpp:19-23)."""
    i = np.arange(n, dtype=np.float64)
    return scale * np.sin(0.37 * i + 0.5)


def prolong3d(coarse: np.ndarray, nc: int) -> np.ndarray:
    """B1 coarse start, experiments/c1/starts/materialize_coarse_start.cpp:60-167:
    fine (nf = 2 nc + 1)^3 from coarse nc^3; per axis an odd fine index takes
    coarse (f-1)/2 with weight 1, an even f = 2I takes I-1 (I >= 1) and I
    (I < nc) with 1/2; the present terms summed in (i, j, k) order."""
    nf = 2 * nc + 1
    c = np.asarray(coarse, np.float64).reshape(nc, nc, nc)

    def terms(f):
        if f % 2 == 1:
            return [((f - 1) // 2, 1.0)]
        r = f // 2
        out = []
        if r >= 1:
            out.append((r - 1, 0.5))
        if r < nc:
            out.append((r, 0.5))
        return out

    T = [terms(f) for f in range(nf)]
    fine = np.zeros((nf, nf, nf))
    for i in range(nf):
        for j in range(nf):
            for k in range(nf):
                s = 0.0
                for a, wa in T[i]:
                    for b, wb in T[j]:
                        for cc, wc in T[k]:
                            s += wa * wb * wc * c[a, b, cc]
                fine[i, j, k] = s
    return fine.ravel()
