/*
 * numa_hip.h - C ABI of the MI355X (gfx950) backend of the matrix-free
 * nonlinear-solver path: Bratu3D residual and directional derivatives,
 * BLAS-1 / Arnoldi reductions, geometric multigrid, (F)GMRES and the
 * Newton / inexact-Chebyshev outer loops, plus the dense Chandrasekhar
 * H-equation operator.
 *
 * Every entry point returns an int status (NUMA_OK = 0).  On failure the
 * message is available from numa_hip_last_error() (thread local).  Numerical
 * outcomes (non-convergence, breakdown) are reported in result structs and
 * never as an error status, as in the reference.
 *
 * Pointer arguments named *_dev are device (HBM) addresses; *_host are host
 * addresses.  All device work of a context is issued on the context's stream.
 * Handles are not thread safe: one handle per host thread.
 *
 * References below are to kapumota/chebyshev-nonlinear-lab, include/numa/.
 */
#ifndef NUMA_HIP_H
#define NUMA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    NUMA_OK = 0,
    NUMA_ERR_ARG = 1,      /* invalid argument (reference: std::invalid_argument) */
    NUMA_ERR_HIP = 2,      /* HIP runtime error                                   */
    NUMA_ERR_NOMEM = 3,    /* device allocation failed                            */
    NUMA_ERR_OVERFLOW = 4, /* size not representable (reference: std::overflow_error) */
    NUMA_ERR_STATE = 5     /* handle used in a state that does not allow the call */
};

/* ---- enumerations used in the POD configs ------------------------------ */
enum { NUMA_METHOD_NEWTON = 0, NUMA_METHOD_CHEBYSHEV = 1, NUMA_METHOD_TF2018 = 2 };
enum {
    NUMA_PRECOND_NONE = 0,
    NUMA_PRECOND_JACOBI = 1, /* BratuJacobi, bratu3d.hpp:104-120            */
    NUMA_PRECOND_GMG = 2,    /* geometric multigrid V-cycle, see numa_mg_config */
    NUMA_PRECOND_LEARNED = 3 /* conv-net preconditioner with V-cycle fallback
                                (no reference counterpart; needs flexible = 1) */
};
enum { NUMA_PRECISION_FP64 = 0, NUMA_PRECISION_MIXED = 1 };
enum { NUMA_DERIV_FD = 0, NUMA_DERIV_ANALYTIC = 1 };
enum { NUMA_FORCING_EW = 0, NUMA_FORCING_CONSTANT = 1 };
enum {
    NUMA_SMOOTH_RBGS = 0,     /* red-black Gauss-Seidel, poisson_mg.hpp:99-126 */
    NUMA_SMOOTH_WJACOBI = 1,  /* weighted Jacobi                               */
    NUMA_SMOOTH_CHEBYSHEV = 2 /* Chebyshev polynomial of D^-1 A                */
};

/* ---- POD parameter / result structs ------------------------------------ */
/* All fields are 8 bytes wide so the layout is identical for C, C++,
 * ctypes and any FFI. */

/* Mirrors numa::op::KrylovParams (methods/krylov_newton.hpp:24-38) field for
 * field, with the additions after `recycle`. */
typedef struct numa_krylov_params {
    double rtol_nl;          /* ||F|| <= rtol_nl * ||F(x0)||           */
    int64_t max_nl;
    double eta_init;
    double eta_max;
    double gamma;
    double alpha;
    double eta2_factor;
    double eta2_max;
    int64_t gmres_restart;
    int64_t gmres_max_iters;
    double armijo;
    double min_lambda;
    int64_t recycle;
    /* additions (no reference counterpart) */
    int64_t forcing;         /* NUMA_FORCING_EW (reference) or _CONSTANT (eta = eta_init) */
    /* Final-solve safeguard (0 = off, the reference): once ||F|| <= tight_res
     * ||F(x0)||, the forcing term is capped at tight_eta, so the last Newton
     * step overshoots the tolerance and the final iterate no longer depends
     * on the path by which it got there. */
    double tight_res;
    double tight_eta;
} numa_krylov_params;

/* Mirrors numa::op::KrylovResult (krylov_newton.hpp:40-49) without the
 * solution vector (returned in place), plus the Timers fields of
 * operators/timing.hpp:16-24 measured with HIP events on the stream. */
typedef struct numa_krylov_result {
    int64_t converged;
    int64_t iters;
    int64_t n_F;
    int64_t n_JVP;
    int64_t n_D2;
    int64_t n_lin_iters;
    int64_t n_lin_failures;
    int64_t n_prec;
    double res0;
    double res;
    double t_total;          /* seconds, wall clock of the whole solve        */
} numa_krylov_result;

/* Mirrors numa::op::GmresResult (linalg/gmres.hpp:14-21). */
typedef struct numa_gmres_result {
    int64_t iters;
    int64_t matvecs;
    int64_t precs;
    double rel_res;
    int64_t converged;
    int64_t breakdown;
} numa_gmres_result;

/* Multigrid V-cycle configuration.  The reference PoissonMG
 * (linalg/poisson_mg.hpp:26-64) is smoother = RBGS, pre = post = 2,
 * even_coarsening = 0, max_levels = 0, min_coarse_n = 1, coarse_iters = 30. */
typedef struct numa_mg_config {
    int64_t smoother;        /* NUMA_SMOOTH_*                                      */
    int64_t pre;             /* pre-smoothing sweeps                               */
    int64_t post;            /* post-smoothing sweeps                              */
    int64_t degree;          /* Chebyshev smoother degree (matvecs per sweep)      */
    int64_t max_levels;      /* 0 = no limit                                       */
    int64_t min_coarse_n;    /* stop coarsening once n <= min_coarse_n             */
    int64_t coarse_iters;    /* coarsest level: RB-GS sweeps or Chebyshev degree;
                                < 0: exact solve by sine transforms (n <= 16)      */
    int64_t even_coarsening; /* 1: coarsen any n to n/2; 0: only odd n (reference) */
    double omega;            /* weighted-Jacobi damping                            */
    double cheb_ratio;       /* Chebyshev smoothing interval [lmax/ratio, lmax]    */
    int64_t fuse;            /* NUMA_FUSE_*: smoothing steps per kernel launch     */
    int64_t ichunk;          /* planes per workgroup of fused launches (0 = 32)    */
    int64_t rows;            /* rows per thread of fused launches (region 64 x 16
                                rows): 0 = 2 (FP32) / 1 (FP64), or twice that;
                                -1 / -2 / -3: box form (<= 6 steps per launch),
                                one workgroup per 8x8x16 / 4x8x16 / 4x4x16 (i x j
                                x k) output nodes, ichunk unused               */
} numa_mg_config;

/* Kernel fusion of the polynomial (weighted-Jacobi / Chebyshev) smoothers.
 * Results are bit for bit the same for every choice. */
enum {
    NUMA_FUSE_NONE = 0,  /* one launch per step                                   */
    NUMA_FUSE_SWEEP = 1, /* one launch per sweep (the degree's steps)             */
    NUMA_FUSE_PHASE = 2, /* one launch for all pre (post) smoothing steps, <= 6   */
    NUMA_FUSE_AUTO = 3   /* per level, the fastest of the above and of the plane
                            chunks {16, 32, 64} and rows per thread, timed on the
                            device at the first
                            V-cycle and cached in a JSON file
                            ($NUMA_HIP_TUNE_CACHE, default
                            ~/.cache/numa_hip/tune.json)                          */
};

/* Solver configuration: what the reference selects by template arguments and
 * provider types, plus the precision / flexibility additions. */
typedef struct numa_solver_config {
    int64_t method;          /* NUMA_METHOD_*                                   */
    int64_t precond;         /* NUMA_PRECOND_*                                  */
    int64_t precision;       /* NUMA_PRECISION_*                                */
    int64_t flexible;        /* 0: GMRES (gmres.hpp:61-223), 1: FGMRES          */
    int64_t jv_mode;         /* NUMA_DERIV_FD (jvp_fd) or _ANALYTIC (Bratu3D::jvp) */
    int64_t d2_mode;         /* NUMA_DERIV_FD (d2_fd) or _ANALYTIC (Bratu3D::d2)   */
    numa_mg_config mg;
} numa_solver_config;

/* ---- opaque handles ---------------------------------------------------- */
typedef struct numa_hip_ctx numa_hip_ctx;
typedef struct numa_hip_mg numa_hip_mg;
typedef struct numa_hip_solver numa_hip_solver;

/* ---- context and memory ------------------------------------------------ */
const char* numa_hip_last_error(void);
int numa_hip_ctx_create(int device, numa_hip_ctx** out);
/* Fails with NUMA_ERR_STATE while multigrid or solver handles created on the
 * context are alive: destroy those first. */
int numa_hip_ctx_destroy(numa_hip_ctx* ctx);
int numa_hip_ctx_sync(numa_hip_ctx* ctx);
int numa_hip_malloc(numa_hip_ctx* ctx, size_t bytes, void** out_dev);
int numa_hip_free(numa_hip_ctx* ctx, void* dev);
int numa_hip_memcpy_h2d(numa_hip_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
int numa_hip_memcpy_d2h(numa_hip_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);

/* ---- Bratu3D operator (problems/bratu3d.hpp) ---------------------------- */
/* F(u) = (n+1)^2 (6u - sum of 6 neighbours) - lambda e^u, zero Dirichlet,
 * index (i*n+j)*n+k.  Replaces Bratu3D::eval (bratu3d.hpp:21-31, 73-95).
 * If norm_host is not NULL it receives ||F||_2 (the call then synchronises). */
int numa_hip_bratu_eval(numa_hip_ctx* ctx, int64_t n, double lambda,
                        const double* u_dev, double* F_dev, double* norm_host);
/* Analytic J v.  Replaces Bratu3D::jvp (bratu3d.hpp:33-43). */
int numa_hip_bratu_jvp(numa_hip_ctx* ctx, int64_t n, double lambda,
                       const double* u_dev, const double* v_dev, double* out_dev);
/* Analytic F''[v,v] = -lambda e^u v^2.  Replaces Bratu3D::d2 (bratu3d.hpp:46-56). */
int numa_hip_bratu_d2(numa_hip_ctx* ctx, int64_t n, double lambda,
                      const double* u_dev, const double* v_dev, double* out_dev);
/* Jacobi diagonal 6(n+1)^2 - lambda e^u.  Replaces Bratu3D::jacobi_diag
 * (bratu3d.hpp:59-69). */
int numa_hip_bratu_jacobi_diag(numa_hip_ctx* ctx, int64_t n, double lambda,
                               const double* u_dev, double* d_dev);
/* Forward-difference J v = (F(x+hv) - F0)/h, h = sqrt(eps)(1+||x||)/||v||,
 * out = 0 if ||v|| = 0, with F0 = F(x) supplied by the caller (the reference
 * recomputes it: operators/operator.hpp:101-114).  FP64 throughout. */
int numa_hip_bratu_jvp_fd(numa_hip_ctx* ctx, int64_t n, double lambda,
                          const double* x_dev, const double* F0_dev,
                          const double* v_dev, double* out_dev);
/* Mixed form of the above: x, F0, F(x+hv) and the subtraction in FP64,
 * v and out in FP32. */
int numa_hip_bratu_jvp_fd_mixed(numa_hip_ctx* ctx, int64_t n, double lambda,
                                const double* x_dev, const double* F0_dev,
                                const float* v_dev, float* out_dev);
/* The mixed solver's J v: the forward-difference quotient of
 * numa_hip_bratu_jvp_fd (operator.hpp:101-114) in the rearranged form
 *   out = h2inv (6v - sum nbrs) - E (e^{hv} - 1)/h,  E = lambda e^x,
 * h = sqrt(eps)(1+||x||)/||v|| (out = 0 if ||v|| = 0): the Laplacian is
 * linear, so only the exponential is differenced, without the rounding of
 * x + hv and of F(x+hv) - F(x).  x in FP64; v, out and E (written here,
 * n^3 floats, the per-Newton-point data the solver keeps) in FP32; the
 * quotient in FP64. */
int numa_hip_bratu_jvp_fd_rearranged(numa_hip_ctx* ctx, int64_t n, double lambda,
                                     const double* x_dev, const float* v_dev, float* E_dev,
                                     float* out_dev);
/* No counterpart in the reference: numa_hip_bratu_jvp_fd without a context,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * the same result bit for bit.  It allocates nothing: ||x||, ||v|| and their
 * reduction partials live in a file-scope device array of the library
 * (4104 doubles, 32.8 KB per device, held until the process ends).  The
 * kernels run on `stream` (a hipStream_t of `device`, NULL = the null
 * stream), which becomes the current device.  n <= 2048, as above.  Calls
 * that overlap on one device share that array and race: serialise them (one
 * stream, or events). */
int numa_hip_bratu_jvp_fd_noalloc(int device, void* stream, int64_t n, double lambda,
                                  const double* x_dev, const double* F0_dev,
                                  const double* v_dev, double* out_dev);
/* Central-difference F''(x)[v,v] = (F(x+hv) - 2F0 + F(x-hv))/h^2,
 * h = eps^(1/4)(1+||x||)/||v||.  Replaces d2_fd (operator.hpp:118-134). */
int numa_hip_bratu_d2_fd(numa_hip_ctx* ctx, int64_t n, double lambda,
                         const double* x_dev, const double* F0_dev,
                         const double* v_dev, double* out_dev);

/* ---- BLAS-1 (operators/operator.hpp:19-41) ------------------------------ */
/* Deterministic fixed-order two-stage reductions, FP64 accumulation. */
int numa_hip_dot(numa_hip_ctx* ctx, int64_t n, const double* a_dev,
                 const double* b_dev, double* result_host);
int numa_hip_dot_f32(numa_hip_ctx* ctx, int64_t n, const float* a_dev,
                     const float* b_dev, double* result_host);
int numa_hip_axpy(numa_hip_ctx* ctx, int64_t n, double a, const double* x_dev,
                  double* y_dev);
/* out = a x (out may be x; a = 1 copies).  operator.hpp:19-41 has no scal; it
 * is the copy / scale the Krylov layer of the C++ drop-in needs. */
int numa_hip_scal(numa_hip_ctx* ctx, int64_t n, double a, const double* x_dev, double* out_dev);

/* ---- geometric multigrid (linalg/poisson_mg.hpp) ------------------------ */
/* Builds the level hierarchy of -Lap_h on n^3 interior nodes.  Replaces
 * PoissonMG::PoissonMG (poisson_mg.hpp:26-64); precision selects FP64 or FP32
 * level storage. */
int numa_hip_mg_create(numa_hip_ctx* ctx, int64_t n, const numa_mg_config* cfg,
                       int64_t precision, numa_hip_mg** out);
int numa_hip_mg_levels(const numa_hip_mg* mg, int64_t* levels, int64_t* sizes_host);
/* z = one V-cycle applied to r (zero initial guess).  Replaces
 * PoissonMG::operator() (poisson_mg.hpp:69-80).  r and z are double* for
 * NUMA_PRECISION_FP64 and float* for NUMA_PRECISION_MIXED. */
int numa_hip_mg_apply(numa_hip_mg* mg, const void* r_dev, void* z_dev);
/* numa_hip_mg_apply and *norm_host = ||z||_2 (FP64 sum of squares of the
 * stored z), taken by the V-cycle's last launch when that is a fused
 * smoothing launch.  The finite-difference J v of operators/operator.hpp:
 * 101-114 needs ||z|| of the preconditioned vector; the solvers use this
 * form.  No counterpart in the reference. */
int numa_hip_mg_apply_norm(numa_hip_mg* mg, const void* r_dev, void* z_dev, double* norm_host);
int numa_hip_mg_destroy(numa_hip_mg* mg);
/* Per level: the smoother fusion in use (NUMA_FUSE_*; NUMA_FUSE_AUTO until the
 * first V-cycle has tuned it), its plane chunk and rows per thread (the
 * fused region is 64 x 16 rows).  Arrays of levels entries. */
int numa_hip_mg_fusion(const numa_hip_mg* mg, int64_t* mode_host, int64_t* ichunk_host,
                       int64_t* rows_host);
/* Residual and restriction of every level in one launch (ich > 0 coarse
 * planes per workgroup; ich < 0 the default chunk, the state after create) or
 * as two launches (ich = 0).  The values are the same bit for bit either way.
 * Tuning / testing knob; no counterpart in the reference. */
int numa_hip_mg_set_resid_restrict(numa_hip_mg* mg, int64_t ich);

/* ---- device-resident solvers on Bratu3D ---------------------------------- */
int numa_hip_bratu_solver_create(numa_hip_ctx* ctx, int64_t n, double lambda,
                                 const numa_solver_config* cfg,
                                 int64_t gmres_restart,
                                 numa_hip_solver** out);
int numa_hip_solver_destroy(numa_hip_solver* s);
/* Solves J(x0) z = b with restarted right-preconditioned GMRES, or FGMRES if
 * the solver was created flexible.  Replaces numa::op::gmres
 * (linalg/gmres.hpp:61-65) with apply = J(x0)v, precond = the configured
 * preconditioner.  x0, b and z are FP64 device vectors. */
int numa_hip_gmres(numa_hip_solver* s, const double* x0_dev, const double* b_dev,
                   double* z_dev, double rtol, int64_t restart, int64_t max_iters,
                   int64_t zero_guess, numa_gmres_result* out);
/* numa_hip_gmres with the rest of gmres.hpp:61-65: r_init_dev (may be NULL)
 * is the initial residual b - J x (no J v spent on it), save != 0 keeps the
 * last Arnoldi cycle (KrylovSpace, gmres.hpp:25-35) in the solver for
 * numa_hip_gmres_space / numa_hip_gmres_recycled. */
int numa_hip_gmres_ex(numa_hip_solver* s, const double* x0_dev, const double* b_dev,
                      double* z_dev, double rtol, int64_t restart, int64_t max_iters,
                      int64_t zero_guess, const double* r_init_dev, int64_t save,
                      numa_gmres_result* out);
/* The saved relation A M^-1 V_k = V_k+1 Hbar: *k, Hbar ((k+1) x k, row major,
 * may be NULL) and, if V_dev is not NULL, V_0..V_k into the k+1 FP64 device
 * vectors it points to. */
int numa_hip_gmres_space(numa_hip_solver* s, int64_t* k, double* hbar_host, double* const* V_dev);
/* gmres_recycled (gmres.hpp:229-323) on J(x0): b projected on the relation
 * (k, Hbar, V_dev[0..k]), x0 = M^-1 V_k y, then GMRES from that residual; with
 * V_dev NULL the relation saved by the last numa_hip_gmres_ex(save = 1).
 * save != 0 keeps the last cycle of that GMRES in turn. */
int numa_hip_gmres_recycled(numa_hip_solver* s, const double* x0_dev, const double* b_dev,
                            double* z_dev, double rtol, int64_t restart, int64_t max_iters,
                            int64_t k, const double* hbar_host, const double* const* V_dev,
                            int64_t save, numa_gmres_result* out);
/* Newton-Krylov or inexact Chebyshev from x (device, FP64, overwritten by
 * the final iterate).  Replaces numa::op::newton_krylov /
 * chebyshev_inexact (methods/krylov_newton.hpp:69-189). */
int numa_hip_solve(numa_hip_solver* s, const numa_krylov_params* p,
                   double* x_dev, numa_krylov_result* out);
/* Same with the iterate in host memory (staged through the device). */
int numa_hip_solve_host(numa_hip_solver* s, const numa_krylov_params* p,
                        double* x_host, numa_krylov_result* out);

/* ---- spectral estimator and continuation (C1 fold calibration) ---------- */
/* Mirrors numa::analysis::EigenEstimate (analysis/spectral.hpp:22-32); the
 * vector is returned separately. */
typedef struct numa_eigen_result {
    int64_t converged;
    double eigenvalue;
    double residual_rel;
    int64_t outer_iterations;
    int64_t linear_iterations;
    int64_t linear_matvecs;
    int64_t preconditioner_applications;
    int64_t linear_failures;
} numa_eigen_result;

/* Arguments of inverse_smallest_eigenpair (analysis/spectral.hpp:34-42):
 * defaults 1e-7, 1e-10, 30, 100, 2000. */
typedef struct numa_eigen_params {
    double eigen_rtol;
    double linear_rtol;
    int64_t max_outer;
    int64_t restart;
    int64_t max_linear;
} numa_eigen_params;

/* Smallest eigenpair of the Bratu Jacobian J(x) by inverse iteration with
 * GMRES preconditioned by the solver's preconditioner
 * (inverse_smallest_eigenpair, analysis/spectral.hpp:34-115: start
 * q_i = 1 + 0.1 sin(0.6180339887498948 (i+1)), Rayleigh quotient, residual
 * ||J q - theta q|| / max(1, |theta|)).  J v is the solver's jv_mode in FP64;
 * the solver must be FP64 with a basis of at least p->restart vectors.
 * x: device FP64; vec_dev (may be NULL): the last iterate q. */
int numa_hip_eigen_smallest(numa_hip_solver* s, const double* x_dev, const numa_eigen_params* p,
                            numa_eigen_result* out, double* vec_dev);

/* Continuation arguments of apps/calibrate_b6.cpp:30-40 (defaults: 0, 9.8,
 * 0.25, 1e-4, 5e-2, 1e-7, 1e-10). */
typedef struct numa_continuation_params {
    double lambda_start;
    double lambda_stop;
    double step;
    double min_step;
    double spectral_stop;
    double eigen_rtol;
    double linear_rtol;
} numa_continuation_params;

/* One accepted point (the CSV row of calibrate_b6.cpp:128-137). */
typedef struct numa_continuation_point {
    double lambda;
    double spectral_gap;
    double eigen_residual;
    int64_t continuation_iters;
    int64_t n_F;
    int64_t n_JVP;
    int64_t n_prec;
} numa_continuation_point;

/* Natural-parameter continuation of the lower Bratu branch with a spectral
 * stop (apps/calibrate_b6.cpp:173-298): Newton-Krylov corrector from the
 * previous point (nk: the corrector's parameters, calibrate_b6.cpp:181-187:
 * rtol_nl 1e-11, max_nl 100, eta 1e-8..1e-6, GMRES(100), 2000), step halved
 * on corrector failure down to min_step, grown by 1.5 after easy points, the
 * smallest eigenvalue of J at every accepted point (30 outer, GMRES(100),
 * 2000), stop when it drops to spectral_stop.  cfg: the solver (FP64;
 * calibrate_b6 uses the analytic J v and PoissonMG(n), i.e. NUMA_PRECOND_GMG
 * with the reference multigrid).  Writes up to max_points points, *count,
 * *status (0 ok; 3 corrector failed below min_step; 4 spectral estimate
 * failed, the app's exit codes), *fold_proxy (the gap^2 linear fit over the
 * last 4 points, calibrate_b6.cpp:139-171; NaN if not admissible) and, if
 * x_out_dev is not NULL, the last accepted solution. */
int numa_hip_continuation(numa_hip_ctx* ctx, int64_t n, const numa_solver_config* cfg,
                          const numa_continuation_params* cp, const numa_krylov_params* nk,
                          numa_continuation_point* points, int64_t max_points, int64_t* count,
                          int64_t* status, double* fold_proxy, double* x_out_dev);

/* ---- start vectors of the C1 campaign ------------------------------------- */
/* One value per line with 17 significant digits (materialize_coarse_start.cpp:
 * 18-43 writes with std::setprecision(max_digits10)). */
int numa_hip_start_write(const char* path, const double* x_host, int64_t count);
/* Whitespace-separated values to the end of the file, all finite, exactly
 * `expected` of them, else NUMA_ERR_ARG (load_vector, run_campaign.cpp:287-302). */
int numa_hip_start_read(const char* path, int64_t expected, double* x_host);
/* SHA-256 (FIPS 180-4) as 64 lower-case hex digits + NUL in hex_out[65]: the
 * sha256 columns of the frozen start manifests (experiments/c1/starts/frozen). */
int numa_hip_sha256(const void* data, int64_t bytes, char* hex_out);
int numa_hip_sha256_file(const char* path, char* hex_out);
/* Perturbed start out = ref + rho ||ref|| d / ||d||, d_i ~ N(0, 1) from
 * std::mt19937_64(seed * 7919 + 13) and std::normal_distribution<double>
 * (perturb, run_campaign.cpp:304-327; norms summed in index order).  Host. */
int numa_hip_perturb_host(const double* ref_host, int64_t count, double rho, uint64_t seed,
                          double* out_host);
/* B1 coarse start: fine = P coarse on the device, fine size 2 nc + 1
 * (prolong3d, experiments/c1/starts/materialize_coarse_start.cpp:60-167). */
int numa_hip_prolong3d(numa_hip_ctx* ctx, const double* coarse_dev, int64_t nc,
                       double* fine_dev);

/* ---- Chandrasekhar H-equation, batch of parameters c ---------------------- */
/* problems/chandrasekhar_op.hpp:31-136 for `batch` values of c at once:
 * (A x)_i = c sum_j 0.5 w_j mu_i/(mu_i+mu_j) x_j with rule 0 = Gauss-Legendre
 * on [0,1] (functions/gauss_legendre_dynamic.hpp:50-114) or rule 1 = midpoint
 * mu_i = (i-1/2)/N, w = 1/N; form 0: F = 1 + x(Ax) - x (reference), form 1:
 * F = x - 1/(1 - Ax).  Vectors are instance-major: x[b*n + i].  The contraction
 * runs on MFMA: FP64 for F; for J v FP64 (precision FP64) or split-BF16 with
 * FP32 accumulation (precision MIXED).  The handle stores the kernel matrix
 * (padded N^2 doubles, and for MIXED N^2 bf16 pairs) when that takes at most
 * 16 GiB and the device can allocate it; NUMA_HIP_CHANDRA_STORE_K=0 in the environment at creation keeps the
 * kernels that generate it per use (the results are bit-identical).  With the stored matrix, the
 * lockstep solve contracts only the instances still iterating (packed blocks of columns, 64 MB more
 * device memory at N=8192, B=256 for MIXED, 32 MB for FP64) and runs its per-instance Arnoldi
 * kernels only for them; NUMA_HIP_CHANDRA_PACK=0 at creation computes every instance every time
 * (same results, bit for bit). */
typedef struct numa_hip_chandra numa_hip_chandra;
int numa_hip_chandra_create(numa_hip_ctx* ctx, int64_t n, int64_t batch, const double* c_host,
                            int64_t rule, int64_t form, int64_t precision,
                            int64_t gmres_restart, numa_hip_chandra** out);
int numa_hip_chandra_destroy(numa_hip_chandra* h);
/* F for every instance (ChandrasekharOp::eval, chandrasekhar_op.hpp:56-66);
 * norms_host (batch doubles, may be NULL) receives ||F_b||. */
int numa_hip_chandra_eval(numa_hip_chandra* h, const double* x_dev, double* F_dev,
                          double* norms_host);
/* J v for every instance (ChandrasekharOp::jvp, chandrasekhar_op.hpp:68-79). */
int numa_hip_chandra_jvp(numa_hip_chandra* h, const double* x_dev, const double* v_dev,
                         double* out_dev);
/* F''(x)[v,v] for every instance (ChandrasekharOp::d2, chandrasekhar_op.hpp:
 * 81-93: 2 v Av; Kelley form: -2 (Av)^2 / (1 - Ax)^3), FP64. */
int numa_hip_chandra_d2(numa_hip_chandra* h, const double* x_dev, const double* v_dev,
                        double* out_dev);
/* The outer method of numa_hip_chandra_solve: NUMA_METHOD_NEWTON (the
 * default, newton_krylov) or NUMA_METHOD_CHEBYSHEV (chebyshev_inexact,
 * krylov_newton.hpp:116-189, with F''[z,z] from numa_hip_chandra_d2). */
int numa_hip_chandra_set_method(numa_hip_chandra* h, int64_t method);
/* Nodes and weights of the rule on [0, 1] (rule 0: gauss_legendre_unit,
 * functions/gauss_legendre_dynamic.hpp:50-114; rule 1: midpoint), host. */
int numa_hip_chandra_quadrature(int64_t n, int64_t rule, double* nodes_host,
                                double* weights_host);
/* newton_krylov (krylov_newton.hpp:69-114, no preconditioner), or the method
 * set above, for every instance; x_dev holds the starting points and receives
 * the iterates; results[batch] receives each instance's KrylovResult. */
int numa_hip_chandra_solve(numa_hip_chandra* h, const numa_krylov_params* p, double* x_dev,
                           numa_krylov_result* results);
/* Average ms of one batched contraction (FP64 or split-BF16) and its useful
 * FLOP count 2 n^2 batch. */
int numa_hip_chandra_gemm_cost(numa_hip_chandra* h, int64_t precision, int64_t reps,
                               double* ms_per_call, double* flops_per_call);

/* ---- learned preconditioner (no reference counterpart) -------------------- */
/* A 3D conv net of six 3x3x3 layers, channels 1-24-24-24-24-24-1, GELU
 * between layers, zero padding, used as z = M(r) = rho/(n+1)^2 net(r/rho),
 * rho = rms(r), in FGMRES with the V-cycle as fallback.  The four 24->24
 * layers run on bf16 MFMA with FP32 accumulation.  Parameters are FP32,
 * flat: per layer W[Cout][Cin][3][3][3] then b[Cout] (63625 values). */
typedef struct numa_hip_net numa_hip_net;

/* Loss of a training pair (x, y) with network output o:
 * CORRECTION: ||o - y||^2 / ||y||^2;
 * RESIDUAL:   ||A7 (o - y)||^2 / ||A7 y||^2, A7 u = 6u - (sum of the six
 *             neighbours), zero outside: the error in the operator's residual. */
enum { NUMA_NET_LOSS_CORRECTION = 0, NUMA_NET_LOSS_RESIDUAL = 1 };

/* Training on pairs generated by the solver: mixed-precision Newton-FGMRES
 * with the V-cycle `mg` on random instances (n from `sizes`, lambda uniform,
 * u0 = noise * U[0,1)); each preconditioner call (r, z) is one pair, with
 * input r/rho and target (n+1)^2 z/rho.  Adam, one step per `batch` pairs. */
typedef struct numa_net_train_config {
    int64_t pairs;            /* training pairs to consume                       */
    int64_t batch;            /* pairs per optimiser step (1..64)                */
    int64_t sizes[4];         /* grid sizes drawn uniformly; 0 = unused slot     */
    double lambda_lo;         /* lambda ~ U[lambda_lo, lambda_hi]                */
    double lambda_hi;
    double noise;             /* u0 = noise * U[0,1) per grid point              */
    double lr;                /* Adam step size                                  */
    int64_t seed;             /* mt19937_64; U[0,1) = (u64 >> 11) * 2^-53        */
    int64_t gmres_restart;    /* FGMRES restart of the generating solves         */
    numa_mg_config mg;        /* V-cycle whose corrections are the targets       */
    numa_krylov_params solve; /* Newton parameters of the generating solves      */
    int64_t loss;             /* NUMA_NET_LOSS_*                                  */
    double lr_end;            /* the step size decays linearly from lr to lr_end
                                 over the pairs (0: lr/10)                       */
} numa_net_train_config;

typedef struct numa_net_train_stats {
    int64_t pairs;      /* pairs consumed                              */
    int64_t steps;      /* optimiser steps                             */
    int64_t solves;     /* generating solves                           */
    double loss_first;  /* mean loss of the first step                 */
    double loss_last;   /* mean loss of the last step                  */
    double seconds;     /* wall time                                   */
} numa_net_train_stats;

int numa_hip_net_create(numa_hip_ctx* ctx, int64_t seed, numa_hip_net** out);
/* The network stays alive while a solver uses it. */
int numa_hip_net_destroy(numa_hip_net* net);
int numa_hip_net_param_count(int64_t* count);
int numa_hip_net_get_params(numa_hip_net* net, float* host);
int numa_hip_net_set_params(numa_hip_net* net, const float* host);
/* out = net(x) for `batch` grids of n^3 points (FP32, device, grid-major). */
int numa_hip_net_forward(numa_hip_net* net, int64_t n, int64_t batch, const float* x_dev,
                         float* out_dev);
/* Mean loss over the batch (NUMA_NET_LOSS_*) of net(x_b) against y_b, and its
 * gradient (host, 63625 floats, parameter layout); no update. */
int numa_hip_net_loss_grad(numa_hip_net* net, int64_t n, int64_t batch, const float* x_dev,
                           const float* y_dev, int64_t loss_kind, double* loss_host,
                           float* grad_host);
/* One Adam step on pairs (r_b, z_b) scaled as in training. */
int numa_hip_net_train_step(numa_hip_net* net, int64_t n, int64_t batch, const float* r_dev,
                            const float* z_dev, int64_t loss_kind, double lr,
                            double* loss_host);
/* z = M(r) on one n^3 grid. */
int numa_hip_net_apply(numa_hip_net* net, int64_t n, const float* r_dev, float* z_dev);
int numa_hip_net_train_solver(numa_hip_net* net, const numa_net_train_config* cfg,
                              double* loss_hist_host, int64_t hist_len,
                              numa_net_train_stats* stats);
/* Average device ms of one inference (what = 0) or one training step of
 * `batch` grids (what = 1) at n^3. */
int numa_hip_net_cost(numa_hip_net* net, int64_t n, int64_t what, int64_t batch, int64_t reps,
                      double* ms_per_call);
/* Attaches the network to a NUMA_PRECOND_LEARNED solver (NULL detaches).  In
 * each linear solve the network serves the FGMRES steps until one reduces the
 * residual estimate by less than fallback_ratio (|g_j+1| > ratio |g_j|); the
 * V-cycle serves the rest of that solve. */
int numa_hip_solver_set_net(numa_hip_solver* s, numa_hip_net* net, double fallback_ratio);
/* The solver's multigrid levels and, per level, the smoother fusion, plane
 * chunk and rows (as numa_hip_mg_fusion); the arrays may be NULL to query the
 * count. */
int numa_hip_solver_mg_fusion(const numa_hip_solver* s, int64_t* levels, int64_t* mode_host,
                              int64_t* ichunk_host, int64_t* rows_host);
/* stats3 = {network applications, V-cycle applications, fallbacks}. */
int numa_hip_solver_learned_stats(const numa_hip_solver* s, int64_t* stats3);

/* ---- timers (operators/timing.hpp:16-108) ------------------------------- */
/* Mirrors numa::op::Timers: seconds spent in F, J v, F'' and the
 * preconditioner's applications and per-point set-up (Provider::make; here the
 * Jacobi diagonal and the ||x|| of the FD step).  On the device each is the
 * GPU time of the call's kernels (a pair of HIP events on the solver stream),
 * accumulated over solves until reset. */
typedef struct numa_timers {
    double t_F;
    double t_JVP;
    double t_D2;
    double t_prec;
    double t_setup_dynamic;
} numa_timers;
/* enable != 0 times every primitive call of the following solves (Timed /
 * TimedPrecond around the operator and the provider); the Arnoldi steps then
 * run eagerly, without graphs. */
int numa_hip_solver_set_timing(numa_hip_solver* s, int64_t enable);
/* Accumulated timers; reset != 0 zeroes them afterwards. */
int numa_hip_solver_timers(numa_hip_solver* s, numa_timers* out, int64_t reset);

/* tensor_free_curvature_2018 (methods/tf_chebyshev_2018.hpp:31-51): out =
 * [J(x + h d) d - J(x) d] / h with h = clamp(sqrt(eps) ||x|| / ||d||, 2 eps,
 * 0.5) (2 eps if d = 0), J v as the solver's jv_mode (analytic, bratu3d.hpp:
 * 33-43, or the forward difference of operator.hpp:101-114 at each base
 * point).  x, d, out: device FP64 vectors of size n^3, out distinct from x
 * and d; h_used (host, may be NULL) receives h.  The TF-2018 outer method is
 * cfg->method = NUMA_METHOD_TF2018 in numa_hip_solve (tf_chebyshev_2018.hpp:
 * 53-132). */
int numa_hip_tf2018_curvature(numa_hip_solver* s, const double* x_dev, const double* d_dev,
                              double* out_dev, double* h_used);

/* ---- Krylov-step graphs and the batched sweep (no reference counterpart) -- */
/* enable != 0: solve() replays each Arnoldi step (preconditioner, J z, MGS
 * passes) from a hipGraph captured at its first use (per step index and
 * iterate buffer); the iterate is copied into the solver's own buffers for
 * the solve.  Results are identical to the eager path.  Solvers with the
 * learned preconditioner run eagerly. */
int numa_hip_solver_set_graphs(numa_hip_solver* s, int64_t enable);
int numa_hip_solver_graphs_cached(const numa_hip_solver* s, int64_t* count);

/* A sweep over many instances (grid n_i, lambda_i) with one configuration:
 * `workers` streams of the context's device, each with its own solver per
 * grid size (reused across instances), driven by as many host threads; the
 * instances, largest first, go to the next free worker. */
typedef struct numa_hip_sweep numa_hip_sweep;
int numa_hip_sweep_create(numa_hip_ctx* ctx, int64_t workers, const numa_solver_config* cfg,
                          int64_t gmres_restart, numa_hip_sweep** out);
int numa_hip_sweep_destroy(numa_hip_sweep* h);
/* Solves instance i from x_dev[i] (device, FP64, n_i^3, overwritten by the
 * solution); results[i] receives its KrylovResult; *seconds (may be NULL) the
 * wall time of the run.  graphs != 0 replays captured Arnoldi steps. */
int numa_hip_sweep_run(numa_hip_sweep* h, int64_t count, const int64_t* n, const double* lambda,
                       double* const* x_dev, const numa_krylov_params* p, int64_t graphs,
                       numa_krylov_result* results, double* seconds);
/* Seeded instances: n_out[i] from sizes (u64 % nsizes), lambda_out[i] ~
 * U[lambda_lo, lambda_hi], std::mt19937_64(seed), U[0,1) = (u64 >> 11) 2^-53. */
int numa_hip_sweep_instances(int64_t seed, int64_t count, const int64_t* sizes, int64_t nsizes,
                             double lambda_lo, double lambda_hi, int64_t* n_out,
                             double* lambda_out);
/* out[i] = scale * U[0,1) from std::mt19937_64(seed) (host only). */
int numa_hip_uniform_host(int64_t seed, int64_t count, double scale, double* out);

/* Batched solver: up to 256 instances (grid n_i <= 127, lambda_i) solved
 * together by mixed-precision Newton-FGMRES(m) with the V-cycle, every
 * operation one kernel launch over the instance table and every Arnoldi step
 * one launch (a hipGraph when graphs are on) for all instances still
 * iterating.  Each instance follows its own Newton / GMRES decisions
 * (newton_krylov, krylov_newton.hpp:69-114; gmres, gmres.hpp:61-223); its
 * iterates are bit for bit those of a numa_hip_solver with the same
 * configuration and unfused smoothing (cfg->mg.fuse = NUMA_FUSE_NONE).
 * Requirements: method Newton, precision mixed, flexible GMRES, the V-cycle
 * with Chebyshev or weighted-Jacobi smoothing and the exact coarse solve
 * (coarse_iters < 0), FD J v, the same number of levels for every instance,
 * 1 <= gmres_restart <= 32.  No reference counterpart (the reference solves
 * one instance per call). */
typedef struct numa_hip_batch numa_hip_batch;
typedef struct numa_batch_stats {
    int64_t launches;       /* kernel launches outside graphs                     */
    int64_t graph_launches; /* Arnoldi steps replayed from graphs                 */
    int64_t arnoldi_steps;  /* lockstep Arnoldi steps                             */
    int64_t host_syncs;     /* stream synchronisations                            */
    double bytes;           /* algorithmic HBM bytes of the active instances      */
    int64_t tables;         /* argument tables on the device                      */
    int64_t table_bytes;
    int64_t graphs;         /* captured Arnoldi steps                             */
} numa_batch_stats;
int numa_hip_batch_create(numa_hip_ctx* ctx, int64_t count, const int64_t* n,
                          const double* lambda, const numa_solver_config* cfg,
                          int64_t gmres_restart, numa_hip_batch** out);
int numa_hip_batch_destroy(numa_hip_batch* h);
int numa_hip_batch_set_graphs(numa_hip_batch* h, int64_t enable);
/* Smoothing form of multigrid levels 0 .. count-1: 0 one launch per smoothing
 * step; one launch per sweep in the box form, 1..3 = box shape 8x8x16,
 * 4x8x16, 4x4x16 output nodes; or in the plane-marching form, 4 / 5 = 2 rows
 * per thread and 16 / 32 planes per workgroup, 6 / 7 = 4 rows and 16 / 32
 * planes.  Every form gives the same values bit for bit.  Without this call
 * the batch times the forms (and the residual / restriction plane chunk) per
 * level at its first solve (cfg->mg.fuse != NUMA_FUSE_NONE) and keeps the
 * fastest (cached under the level's grid sizes). */
int numa_hip_batch_set_smoothing(numa_hip_batch* h, int64_t count, const int64_t* form);
/* Solves instance i from x_dev[i] (device, FP64, n_i^3, overwritten by the
 * solution) on the context's stream; results[i] its KrylovResult; *seconds
 * (may be NULL) the wall time. */
int numa_hip_batch_solve(numa_hip_batch* h, const numa_krylov_params* p, double* const* x_dev,
                         numa_krylov_result* results, double* seconds);
/* Counters since creation or the last reset (reset != 0 clears them after
 * reading). */
int numa_hip_batch_stats(numa_hip_batch* h, int64_t reset, numa_batch_stats* out);

/* ---- fault injection (tests of the failure paths) ------------------------ */
/* Forces a non-finite value into one primitive call of the solver, from the
 * host, so that the numerical failure guards of the reference (the
 * non-finite Givens denominator and back-substitution pivot, gmres.hpp:
 * 165-169, 185-190; the line search, krylov_newton.hpp:96-104) are exercised
 * on the device: the `call`-th F evaluation (what = NUMA_INJECT_F: element 0
 * of F and ||F|| become NaN) or J v product (what = NUMA_INJECT_JV: element 0
 * of the product and its fused reduction become NaN), counted from this call
 * over the following solves; it fires once.  call = 0 disarms.  While armed
 * the Arnoldi steps run eagerly.  No counterpart in the reference, whose
 * tests force these paths by wrapping the operator (test_operators.cpp:
 * 141-157). */
enum { NUMA_INJECT_F = 0, NUMA_INJECT_JV = 1 };
int numa_hip_solver_inject_nan(numa_hip_solver* s, int64_t what, int64_t call);

/* ---- per-call primitive costs (instrumentation) -------------------------- */
/* The reference freezes per-call costs c_F, c_JVP, c_D2, c_prec of its
 * primitives (experiments/c1/frozen/primitive_costs.csv, measured by
 * scripts/freeze_primitive_costs.py, semantics docs/research/C1_WEQ_PROTOCOL.md)
 * and times them with Timed / TimedPrecond (operators/timing.hpp:37-108).
 * This entry point measures one primitive of a solver on its own stream with
 * HIP events: average milliseconds per call over `reps` calls, and the call's
 * algorithmic HBM bytes. */
enum {
    NUMA_PRIM_F = 0,      /* F(x), FP64, fused ||F||^2                      */
    NUMA_PRIM_JVP = 1,    /* J v stencil kernel with the fused first MGS dot */
    NUMA_PRIM_D2 = 2,     /* F''[v,v]                                        */
    NUMA_PRIM_PREC = 3,   /* one preconditioner application (V-cycle)        */
    NUMA_PRIM_SMOOTH = 4, /* one finest-level smoothing step                 */
    NUMA_PRIM_MGS = 5,    /* one fused MGS pass (w -= h V_i; dot(w, V_i+1))  */
    NUMA_PRIM_COPY = 6    /* FP64 vector copy (the measured HBM peak)        */
};
int numa_hip_primitive_cost(numa_hip_solver* s, int64_t prim, int64_t reps,
                            double* ms_per_call, double* bytes_per_call);
/* No counterpart in the reference (measurement): the device's dense MFMA
 * throughput in TFLOP/s, measured on the context's stream with HIP events,
 * kind 0 = v_mfma_f32_32x32x16_bf16, 1 = v_mfma_f64_16x16x4f64: 4 or 8
 * workgroups of 4 waves per CU, every wave 4 or 8 independent accumulator
 * chains, no memory traffic; the best of the four shapes (each the best of 2
 * launches after a warm-up).  The measured peak beside the published one for
 * the MFMA kernels' fractions. */
int numa_hip_mfma_peak(numa_hip_ctx* ctx, int64_t kind, double* tflops);
/* In-solve timing of the finest-level smoothing sweep (the roofline kernel of
 * the benchmark): with enable != 0, every finest-level fused smoothing launch
 * of the V-cycle in the plane-marching form that reads the iterate (not the
 * first one from zero, nor the one that also sums ||z||) is launched with a
 * start / stop event pair that the dispatch itself stamps
 * (hipExtLaunchKernelGGL: the kernel's own execution), and the solver's
 * Arnoldi steps run eagerly (no graphs).
 * numa_hip_solver_sweep_time synchronises the stream and returns the average
 * milliseconds per such launch since the last call and their count. */
int numa_hip_solver_sweep_events(numa_hip_solver* s, int64_t enable);
int numa_hip_solver_sweep_time(numa_hip_solver* s, double* avg_ms, int64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* NUMA_HIP_H */
