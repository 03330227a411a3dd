/*
 * numa_oracle.h - C ABI of the CPU FP64 oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a CPU restatement of the
 * reference's matrix-free solver path (kapumota/chebyshev-nonlinear-lab,
 * include/numa/...) plus the new components the GPU path adds (even-n
 * multigrid, weighted-Jacobi and Chebyshev smoothers, FGMRES, constant
 * forcing, midpoint / Kelley-form Chandrasekhar).  Only tests/, the smoke()
 * check and bench.py's cpu_baseline leg may load it, and only as the reference
 * the GPU results are compared with or the timed CPU baseline, never as a
 * product path.
 *
 * Parity pins for every function are listed in the header comment of
 * numa_oracle.cpp.
 */
#ifndef NUMA_ORACLE_H
#define NUMA_ORACLE_H

#include <stdint.h>
#include "../include/numa_hip.h" /* shared POD config / result structs only */

#ifdef __cplusplus
extern "C" {
#endif

int orc_num_threads(void);
/* Fault injection: NaN into element 0 of the `at`-th Bratu F evaluation
 * (what = NUMA_INJECT_F) or J v product (NUMA_INJECT_JV) of the following
 * solves of this thread; fires once; at = 0 disarms. */
void orc_set_inject(int64_t what, int64_t at);

/* Bratu3D (problems/bratu3d.hpp) */
void orc_bratu_eval(int64_t n, double lambda, const double* u, double* out);
void orc_bratu_jvp(int64_t n, double lambda, const double* u, const double* v, double* out);
void orc_bratu_d2(int64_t n, double lambda, const double* u, const double* v, double* out);
void orc_bratu_jacobi_diag(int64_t n, double lambda, const double* u, double* d);
void orc_bratu_jvp_fd(int64_t n, double lambda, const double* x, const double* v, double* out);
void orc_bratu_d2_fd(int64_t n, double lambda, const double* x, const double* v, double* out);
void orc_poisson_apply(int64_t n, const double* v, double* out);

/* BLAS-1 (operators/operator.hpp:19-41) */
double orc_dot(int64_t n, const double* a, const double* b);

/* multigrid (linalg/poisson_mg.hpp) */
int64_t orc_mg_levels(int64_t n, const numa_mg_config* cfg, int64_t* sizes);
int orc_mg_apply(int64_t n, const numa_mg_config* cfg, const double* r, double* z);

/* GMRES on the Bratu Jacobian at x0 (linalg/gmres.hpp:61-223); flexible = 1
 * selects FGMRES.  precond / jv_mode / mg come from cfg. */
int orc_bratu_gmres(int64_t n, double lambda, const numa_solver_config* cfg,
                    const double* x0, const double* b, double* x, double rtol,
                    int64_t restart, int64_t max_iters, int64_t zero_guess,
                    numa_gmres_result* out);
/* GMRES on the plain Poisson operator -Lap_h (test_operators.cpp:207-233). */
int orc_poisson_gmres(int64_t n, const numa_mg_config* mg, int64_t use_mg,
                      const double* b, double* x, double rtol, int64_t restart,
                      int64_t max_iters, numa_gmres_result* out);

/* Newton-Krylov / inexact Chebyshev on Bratu3D (methods/krylov_newton.hpp)
 * or tensor-free Chebyshev 2018 (methods/tf_chebyshev_2018.hpp), cfg->method */
int orc_bratu_solve(int64_t n, double lambda, const numa_solver_config* cfg,
                    const numa_krylov_params* p, double* x, numa_krylov_result* out);
/* inverse_smallest_eigenpair (analysis/spectral.hpp:34-115) of J(x) with the
 * preconditioner of cfg; vec (may be NULL) receives the last iterate. */
int orc_bratu_eigen_smallest(int64_t n, double lambda, const numa_solver_config* cfg,
                             const double* x, const numa_eigen_params* p, numa_eigen_result* out,
                             double* vec);
/* calibrate_b6 continuation (apps/calibrate_b6.cpp:139-298). */
int orc_bratu_continuation(int64_t n, const numa_solver_config* cfg,
                           const numa_continuation_params* cp, const numa_krylov_params* nk,
                           numa_continuation_point* points, int64_t max_points, int64_t* count,
                           int64_t* status, double* fold_proxy, double* x_out);
/* tensor_free_curvature_2018 (tf_chebyshev_2018.hpp:31-51) on Bratu3D with
 * the J v of jv_mode (NUMA_DERIV_ANALYTIC or NUMA_DERIV_FD). */
int orc_tf2018_curvature(int64_t n, double lambda, int64_t jv_mode, const double* x,
                         const double* d, double* out, double* h_used);

/* Chandrasekhar H-equation (problems/chandrasekhar_op.hpp and the midpoint /
 * Kelley form).  rule: 0 = Gauss-Legendre on [0,1], 1 = midpoint.
 * form: 0 = F = 1 + x(Ax) - x (reference), 1 = F = x - 1/(1 - Ax). */
int orc_gauss_legendre_unit(int64_t n, double* nodes, double* weights);
int orc_chandra_eval(int64_t n, double c, int64_t rule, int64_t form,
                     const double* x, double* out);
int orc_chandra_jvp(int64_t n, double c, int64_t rule, int64_t form,
                    const double* x, const double* v, double* out);
int orc_chandra_d2(int64_t n, double c, int64_t rule, int64_t form,
                   const double* x, const double* v, double* out);
/* Newton-GMRES (no preconditioner) or Chebyshev; method from cfg->method. */
int orc_chandra_solve(int64_t n, double c, int64_t rule, int64_t form,
                      const numa_solver_config* cfg, const numa_krylov_params* p,
                      double* x, numa_krylov_result* out);

#ifdef __cplusplus
}
#endif
#endif
