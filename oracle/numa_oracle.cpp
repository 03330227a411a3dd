// numa_oracle.cpp - CPU FP64 oracle for the matrix-free Bratu3D / Chandrasekhar
// solver path.
//
// TEST INFRASTRUCTURE ONLY (see numa_oracle.h).  Written from the reference's
// documented behaviour; every function cites the reference file:line it
// restates (paths relative to the reference's include/numa/).
//
// Parity pins (tests/test_oracle_*.py):
//   bratu eval/jvp/d2/jacobi_diag  - bit-level against the reference headers
//       compiled into oracle/_ref (tests/golden fixtures), closed-form
//       Laplacian eigenvectors, F(0) = -lambda, brute-force dense assembly.
//   jvp_fd / d2_fd                 - against _ref, and analytic vs FD as in
//       test_operators.cpp:39-61.
//   dot                            - exact sums of integers.
//   PoissonMG (RB-GS, odd n)       - against _ref V-cycle output; levels(15)==4.
//   even-n GMG, WJ / Chebyshev     - Chebyshev polynomial closed form on a
//       Laplacian eigenvector; V-cycle reduction on eigenvectors; bounded
//       MG-GMRES iteration counts.
//   gmres / fgmres                 - against _ref, residual truth, dense LU
//       brute force, breakdown cases; FGMRES == GMRES for linear M.
//   newton_krylov / chebyshev      - against _ref counts and iterates; the
//       reference-held continuation norms (B1_start_hashes.csv).
//   smallest eigenpair             - against _ref (inverse_smallest_eigenpair)
//       and the closed form 12 (n+1)^2 sin^2(pi / (2(n+1))) at lambda = 0.
//   continuation (calibrate_b6)    - the reference's recorded fold rows
//       (fold_calibration/bratu_n31_coarse.csv): gaps and counts.
//   tf2018 curvature / solve       - against _ref (tensor_free_curvature_2018,
//       tf_chebyshev_2018) counts and iterates; curvature vs analytic F''
//       (test_tf_chebyshev_2018.cpp:19-33); exact quadratic: for F quadratic
//       the curvature is F''[d,d] to rounding.
//   chandrasekhar                  - N=8 golden vector (test_chandrasekhar_op.cpp:106-115),
//       moment identity sum w x = (2/c)(1 - sqrt(1-c)), GL monomial exactness.

#include "numa_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using DVec = std::vector<double>;

// ---------------------------------------------------------------------------
// BLAS-1.  operators/operator.hpp:19-41.  The reference reduces with an
// OpenMP static reduction whose order depends on the thread count; the oracle
// sums fixed chunks in a fixed order so its result does not depend on it.
// ---------------------------------------------------------------------------
constexpr long kChunk = 16384;

double dot(const double* a, const double* b, long n) {
    const long nch = (n + kChunk - 1) / kChunk;
    if (nch <= 1) {
        double s = 0.0;
        for (long i = 0; i < n; ++i) s += a[i] * b[i];
        return s;
    }
    std::vector<double> part(static_cast<size_t>(nch));
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nch; ++c) {
        const long lo = c * kChunk, hi = std::min(n, lo + kChunk);
        double s = 0.0;
        for (long i = lo; i < hi; ++i) s += a[i] * b[i];
        part[static_cast<size_t>(c)] = s;
    }
    double s = 0.0;
    for (double p : part) s += p;
    return s;
}
double dot(const DVec& a, const DVec& b) { return dot(a.data(), b.data(), (long)a.size()); }
double norm2(const DVec& a) { return std::sqrt(dot(a, a)); }

// y += a x   (operator.hpp:32-41)
void axpy(double a, const DVec& x, DVec& y) {
    const long n = (long)x.size();
#pragma omp parallel for schedule(static) if (n > 4096)
    for (long i = 0; i < n; ++i) y[(size_t)i] += a * x[(size_t)i];
}

// ---------------------------------------------------------------------------
// Bratu3D.  problems/bratu3d.hpp:12-100.
// ---------------------------------------------------------------------------
struct Bratu {
    long n;
    double lambda;
    double h2inv;
    Bratu(long n_, double l) : n(n_), lambda(l), h2inv((n_ + 1.0) * (n_ + 1.0)) {}
    long size() const { return n * n * n; }

    // out = (-Lap_h) v, bratu3d.hpp:73-95 (same neighbour order).
    void lap(const double* v, double* out) const {
        const long nn = n;
#pragma omp parallel for collapse(2) schedule(static) if (nn > 16)
        for (long i = 0; i < nn; ++i)
            for (long j = 0; j < nn; ++j)
                for (long k = 0; k < nn; ++k) {
                    const long id = (i * nn + j) * nn + k;
                    double s = 6.0 * v[id];
                    if (i > 0) s -= v[id - nn * nn];
                    if (i + 1 < nn) s -= v[id + nn * nn];
                    if (j > 0) s -= v[id - nn];
                    if (j + 1 < nn) s -= v[id + nn];
                    if (k > 0) s -= v[id - 1];
                    if (k + 1 < nn) s -= v[id + 1];
                    out[id] = h2inv * s;
                }
    }
    // bratu3d.hpp:21-31
    void eval(const double* u, double* out) const {
        lap(u, out);
        const long N = size();
#pragma omp parallel for schedule(static) if (N > 4096)
        for (long i = 0; i < N; ++i) out[i] -= lambda * std::exp(u[i]);
    }
    // bratu3d.hpp:33-43
    void jvp(const double* u, const double* v, double* out) const {
        lap(v, out);
        const long N = size();
#pragma omp parallel for schedule(static) if (N > 4096)
        for (long i = 0; i < N; ++i) out[i] -= lambda * std::exp(u[i]) * v[i];
    }
    // bratu3d.hpp:46-56
    void d2(const double* u, const double* v, double* out) const {
        const long N = size();
#pragma omp parallel for schedule(static) if (N > 4096)
        for (long i = 0; i < N; ++i) out[i] = -lambda * std::exp(u[i]) * v[i] * v[i];
    }
    // bratu3d.hpp:59-69
    void jacobi_diag(const double* u, double* d) const {
        const long N = size();
#pragma omp parallel for schedule(static) if (N > 4096)
        for (long i = 0; i < N; ++i) d[i] = 6.0 * h2inv - lambda * std::exp(u[i]);
    }
};

// Forward-difference J v, operator.hpp:101-114.  F0 = F(x) and ||x|| may be
// supplied (they are what the reference recomputes on every call).
void jvp_fd(const Bratu& f, const double* x, const double* v, double* out,
            const double* F0 = nullptr, double xnorm = -1.0) {
    const long N = f.size();
    const double nv = std::sqrt(dot(v, v, N));
    if (nv == 0.0) {
        std::fill(out, out + N, 0.0);
        return;
    }
    if (xnorm < 0.0) xnorm = std::sqrt(dot(x, x, N));
    const double h = std::sqrt(std::numeric_limits<double>::epsilon()) * (1.0 + xnorm) / nv;
    DVec xp(x, x + N), f1((size_t)N);
    DVec f0;
    for (long i = 0; i < N; ++i) xp[(size_t)i] += h * v[i];
    if (!F0) {
        f0.resize((size_t)N);
        f.eval(x, f0.data());
        F0 = f0.data();
    }
    f.eval(xp.data(), f1.data());
    for (long i = 0; i < N; ++i) out[i] = (f1[(size_t)i] - F0[i]) / h;
}

// Central-difference F''(x)[v,v], operator.hpp:118-134.
void d2_fd(const Bratu& f, const double* x, const double* v, double* out,
           const double* F0 = nullptr, double xnorm = -1.0) {
    const long N = f.size();
    const double nv = std::sqrt(dot(v, v, N));
    if (nv == 0.0) {
        std::fill(out, out + N, 0.0);
        return;
    }
    if (xnorm < 0.0) xnorm = std::sqrt(dot(x, x, N));
    const double h = std::pow(std::numeric_limits<double>::epsilon(), 0.25) * (1.0 + xnorm) / nv;
    DVec xp(x, x + N), xm(x, x + N), fp((size_t)N), fm((size_t)N), f0;
    for (long i = 0; i < N; ++i) xp[(size_t)i] += h * v[i];
    for (long i = 0; i < N; ++i) xm[(size_t)i] += -h * v[i];
    if (!F0) {
        f0.resize((size_t)N);
        f.eval(x, f0.data());
        F0 = f0.data();
    }
    f.eval(xp.data(), fp.data());
    f.eval(xm.data(), fm.data());
    for (long i = 0; i < N; ++i) out[i] = (fp[(size_t)i] - 2.0 * F0[i] + fm[(size_t)i]) / (h * h);
}

// ---------------------------------------------------------------------------
// Geometric multigrid for -Lap_h.  linalg/poisson_mg.hpp:24-320 for the
// RB-GS / odd-n reference hierarchy; the even-n coarsening (n -> n/2, coarse
// node I <-> fine node 2I+1, same transfer operators), weighted Jacobi and
// Chebyshev smoothers and the Chebyshev coarse solve are new.
// ---------------------------------------------------------------------------
struct MG {
    struct Level {
        long n = 0;
        double h2inv = 0.0;
        DVec x, b, r, t;  // t: second buffer for Jacobi-type sweeps / Chebyshev direction
        DVec d;
    };
    numa_mg_config cfg;
    std::vector<Level> levels;

    MG(long n, const numa_mg_config& c) : cfg(c) {
        long m = n;
        double h2inv = (double(n) + 1.0) * (double(n) + 1.0);
        while (true) {
            Level L;
            L.n = m;
            L.h2inv = h2inv;
            const size_t N = (size_t)(m * m * m);
            L.x.assign(N, 0.0);
            L.b.assign(N, 0.0);
            L.r.assign(N, 0.0);
            if (cfg.smoother != NUMA_SMOOTH_RBGS) {
                L.t.assign(N, 0.0);
                L.d.assign(N, 0.0);
            }
            levels.push_back(std::move(L));
            // poisson_mg.hpp:60-62: the reference stops at even m.
            if (m <= 1) break;
            if (!cfg.even_coarsening && (m % 2 == 0)) break;
            if (m <= std::max<int64_t>(1, cfg.min_coarse_n)) break;
            if (cfg.max_levels > 0 && (int64_t)levels.size() >= cfg.max_levels) break;
            m = m / 2;
            h2inv /= 4.0;
        }
    }

    static long idx(long n, long i, long j, long k) { return (i * n + j) * n + k; }

    // Sum of the 6 neighbours (zero outside), in the order of
    // poisson_mg.hpp:114-119.
    static inline double nbr_sum(const double* x, long n, long i, long j, long k, long id) {
        double s = 0.0;
        if (i > 0) s += x[id - n * n];
        if (i + 1 < n) s += x[id + n * n];
        if (j > 0) s += x[id - n];
        if (j + 1 < n) s += x[id + n];
        if (k > 0) s += x[id - 1];
        if (k + 1 < n) s += x[id + 1];
        return s;
    }
    // h2inv (6 x - sum nbrs) in the order of poisson_mg.hpp:135-145.
    static inline double apply_at(const double* x, long n, double h2inv, long i, long j, long k, long id) {
        double s = 6.0 * x[id];
        if (i > 0) s -= x[id - n * n];
        if (i + 1 < n) s -= x[id + n * n];
        if (j > 0) s -= x[id - n];
        if (j + 1 < n) s -= x[id + n];
        if (k > 0) s -= x[id - 1];
        if (k + 1 < n) s -= x[id + 1];
        return h2inv * s;
    }

    // poisson_mg.hpp:99-126.  Within one colour the updates are independent,
    // so the parallel loop gives the sequential result.
    static void smooth_rbgs(Level& L, int order) {
        const long n = L.n;
        const double inv = 1.0 / L.h2inv;
        for (int pass = 0; pass < 2; ++pass) {
            const long color = (order == 0 ? pass : 1 - pass);
#pragma omp parallel for schedule(static) if (n > 16)
            for (long i = 0; i < n; ++i)
                for (long j = 0; j < n; ++j)
                    for (long k = (i + j + color) & 1; k < n; k += 2) {
                        const long id = idx(n, i, j, k);
                        double sum = L.b[(size_t)id] * inv;
                        sum += nbr_sum(L.x.data(), n, i, j, k, id);
                        L.x[(size_t)id] = sum / 6.0;
                    }
        }
    }

    // b - A x at a node with the smoothers' fused multiply-adds:
    // s = fma(6, x, -x_im) - x_ip - x_jm - x_jp - x_km - x_kp (missing
    // neighbours count as 0), res = fma(-h2inv, s, b).
    static inline double res_fma_at(const double* x, const double* b, long n, double h2inv,
                                     long i, long j, long k, long id) {
        double s = std::fma(6.0, x[id], i > 0 ? -x[id - n * n] : -0.0);
        if (i + 1 < n) s -= x[id + n * n];
        if (j > 0) s -= x[id - n];
        if (j + 1 < n) s -= x[id + n];
        if (k > 0) s -= x[id - 1];
        if (k + 1 < n) s -= x[id + 1];
        return std::fma(-h2inv, s, b[id]);
    }

    // One Chebyshev / Jacobi-type step:  d <- c1 d + c2 (b - A x);  x <- x + d.
    // With c1 = 0, c2 = omega/(6 h2inv) it is one weighted-Jacobi sweep.  The
    // smoother (fused) rounds as the device smoothing kernels do: res by
    // res_fma_at, d = fma(c1, d, c2 res); the coarsest-level iteration keeps
    // separate products (as the device's one-workgroup coarse solve).
    static void poly_step(Level& L, double c1, double c2, bool use_d, bool fused = false) {
        const long n = L.n;
        const double* x = L.x.data();
        const double* b = L.b.data();
        double* xn = L.t.data();
        double* d = L.d.data();
#pragma omp parallel for collapse(2) schedule(static) if (n > 16)
        for (long i = 0; i < n; ++i)
            for (long j = 0; j < n; ++j)
                for (long k = 0; k < n; ++k) {
                    const long id = idx(n, i, j, k);
                    double res, dn;
                    if (fused) {
                        res = res_fma_at(x, b, n, L.h2inv, i, j, k, id);
                        dn = use_d ? std::fma(c1, d[id], c2 * res) : c2 * res;
                    } else {
                        res = b[id] - apply_at(x, n, L.h2inv, i, j, k, id);
                        dn = use_d ? c1 * d[id] + c2 * res : c2 * res;
                    }
                    d[id] = dn;
                    xn[id] = x[id] + dn;
                }
        L.x.swap(L.t);
    }

    // Chebyshev iteration on D^-1 A over [alpha, beta], `deg` steps.
    static void chebyshev(Level& L, int deg, double alpha, double beta, bool fused = false) {
        const double theta = 0.5 * (beta + alpha), delta = 0.5 * (beta - alpha);
        const double sigma = theta / delta;
        const double dinv = 1.0 / (6.0 * L.h2inv);
        double rho = 1.0 / sigma;
        poly_step(L, 0.0, dinv / theta, false, fused);
        for (int k = 1; k < deg; ++k) {
            const double rho_new = 1.0 / (2.0 * sigma - rho);
            poly_step(L, rho_new * rho, 2.0 * rho_new / delta * dinv, true, fused);
            rho = rho_new;
        }
    }

    static double lmax(long n) { return 1.0 + std::cos(M_PI / (double(n) + 1.0)); }
    static double lmin(long n) { return 1.0 - std::cos(M_PI / (double(n) + 1.0)); }

    void smooth(Level& L, int order) const {
        switch (cfg.smoother) {
            case NUMA_SMOOTH_RBGS: smooth_rbgs(L, order); break;
            case NUMA_SMOOTH_WJACOBI:
                poly_step(L, 0.0, cfg.omega / (6.0 * L.h2inv), false, true);
                break;
            default: {
                const double b = lmax(L.n);
                chebyshev(L, (int)cfg.degree, b / cfg.cheb_ratio, b, true);
            }
        }
    }

    // poisson_mg.hpp:128-149
    static void compute_residual(Level& L) {
        const long n = L.n;
#pragma omp parallel for collapse(2) schedule(static) if (n > 16)
        for (long i = 0; i < n; ++i)
            for (long j = 0; j < n; ++j)
                for (long k = 0; k < n; ++k) {
                    const long id = idx(n, i, j, k);
                    L.r[(size_t)id] = L.b[(size_t)id] - apply_at(L.x.data(), n, L.h2inv, i, j, k, id);
                }
    }

    // poisson_mg.hpp:152-188 (coarse I <-> fine 2I+1, weights 1/4 1/2 1/4).
    static void restrict_to(const Level& F, Level& C) {
        const long nf = F.n, nc = C.n;
        static const double w[3] = {0.25, 0.5, 0.25};
#pragma omp parallel for collapse(2) schedule(static) if (nc > 8)
        for (long I = 0; I < nc; ++I)
            for (long J = 0; J < nc; ++J)
                for (long K = 0; K < nc; ++K) {
                    double sum = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        const long fi = 2 * I + a;
                        if (fi >= nf) continue;
                        for (int b = 0; b < 3; ++b) {
                            const long fj = 2 * J + b;
                            if (fj >= nf) continue;
                            for (int c = 0; c < 3; ++c) {
                                const long fk = 2 * K + c;
                                if (fk >= nf) continue;
                                sum += w[a] * w[b] * w[c] * F.r[(size_t)idx(nf, fi, fj, fk)];
                            }
                        }
                    }
                    C.b[(size_t)idx(nc, I, J, K)] = sum;
                }
    }

    // poisson_mg.hpp:191-220
    static void interp_terms(long f, long nc, long ci[2], double wi[2], int& cnt) {
        cnt = 0;
        if (f % 2 == 1) {
            ci[0] = (f - 1) / 2;
            wi[0] = 1.0;
            cnt = 1;
            return;
        }
        const long right = f / 2;
        if (right >= 1) {
            ci[cnt] = right - 1;
            wi[cnt] = 0.5;
            ++cnt;
        }
        if (right < nc) {
            ci[cnt] = right;
            wi[cnt] = 0.5;
            ++cnt;
        }
    }

    // poisson_mg.hpp:222-260
    static void prolong_add(const Level& C, Level& F) {
        const long nf = F.n, nc = C.n;
#pragma omp parallel for schedule(static) if (nf > 16)
        for (long i = 0; i < nf; ++i) {
            long ci[2], cj[2], ck[2];
            double wi[2], wj[2], wk[2];
            int ni, nj, nk;
            interp_terms(i, nc, ci, wi, ni);
            for (long j = 0; j < nf; ++j) {
                interp_terms(j, nc, cj, wj, nj);
                for (long k = 0; k < nf; ++k) {
                    interp_terms(k, nc, ck, wk, nk);
                    double sum = 0.0;
                    for (int a = 0; a < ni; ++a)
                        for (int b = 0; b < nj; ++b)
                            for (int c = 0; c < nk; ++c)
                                sum += wi[a] * wj[b] * wk[c] * C.x[(size_t)idx(nc, ci[a], cj[b], ck[c])];
                    F.x[(size_t)idx(nf, i, j, k)] += sum;
                }
            }
        }
    }

    // Exact solve by the discrete sine transform (the GPU's coarse_iters < 0):
    // x = (2/(n+1))^3 S3 diag(1 / (h2inv (c_p + c_q + c_r))) S3 b with
    // S_pk = sin(pi (p+1)(k+1)/(n+1)), c_p = 2 - 2 cos(pi (p+1)/(n+1)).
    static void coarse_dst(Level& L) {
        const long n = L.n, N2 = n * n, N = N2 * n;
        const double th = M_PI / (double)(n + 1);
        std::vector<double> S((size_t)(n * n)), c((size_t)n), u(L.b.begin(), L.b.begin() + N),
            w((size_t)N);
        for (long p = 0; p < n; ++p) {
            c[(size_t)p] = 2.0 - 2.0 * std::cos(th * (double)(p + 1));
            for (long k = 0; k < n; ++k) S[(size_t)(p * n + k)] = std::sin(th * (double)((p + 1) * (k + 1)));
        }
        auto pass = [&](long st, long A, long C) {
            for (long l = 0; l < N2; ++l) {
                const long base = (l / n) * A + (l % n) * C;
                for (long p = 0; p < n; ++p) {
                    double s = 0.0;
                    for (long k = 0; k < n; ++k) s += S[(size_t)(p * n + k)] * u[(size_t)(base + k * st)];
                    w[(size_t)(base + p * st)] = s;
                }
            }
            u.swap(w);
        };
        pass(1, N2, n);
        pass(n, N2, 1);
        pass(N2, n, 1);
        const double f = 8.0 / ((double)(n + 1) * (double)(n + 1) * (double)(n + 1));
        for (long id = 0; id < N; ++id) {
            const long i = id / N2, j = (id / n) % n, k = id % n;
            u[(size_t)id] = f * (u[(size_t)id] / (L.h2inv * (c[(size_t)i] + c[(size_t)j] + c[(size_t)k])));
        }
        pass(1, N2, n);
        pass(n, N2, 1);
        pass(N2, n, 1);
        std::copy(u.begin(), u.end(), L.x.begin());
    }

    void coarse_solve(Level& L) const {
        if (cfg.coarse_iters < 0) {
            coarse_dst(L);
            return;
        }
        if (L.n == 1) {  // poisson_mg.hpp:267-269
            L.x[0] = L.b[0] / (6.0 * L.h2inv);
            return;
        }
        if (cfg.smoother == NUMA_SMOOTH_RBGS) {  // poisson_mg.hpp:270-274
            for (int s = 0; s < (int)cfg.coarse_iters; ++s) smooth_rbgs(L, s % 2);
            return;
        }
        chebyshev(L, (int)cfg.coarse_iters, lmin(L.n), lmax(L.n));
    }

    // poisson_mg.hpp:262-294
    void cycle(size_t l) {
        Level& L = levels[l];
        if (l + 1 == levels.size()) {
            coarse_solve(L);
            return;
        }
        for (int s = 0; s < (int)cfg.pre; ++s) smooth(L, 0);
        compute_residual(L);
        Level& C = levels[l + 1];
        restrict_to(L, C);
        std::fill(C.x.begin(), C.x.end(), 0.0);
        cycle(l + 1);
        prolong_add(C, L);
        for (int s = 0; s < (int)cfg.post; ++s) smooth(L, 1);
    }

    // poisson_mg.hpp:69-80
    void apply(const double* r, double* z) {
        Level& F = levels[0];
        std::copy(r, r + F.x.size(), F.b.begin());
        std::fill(F.x.begin(), F.x.end(), 0.0);
        cycle(0);
        std::copy(F.x.begin(), F.x.end(), z);
    }
};

// ---------------------------------------------------------------------------
// GMRES.  linalg/gmres.hpp:61-223 (same MGS / Givens / back-substitution
// order, same breakdown rules and counters).  flexible = true stores
// Z_j = M^-1 V_j and updates x += Z y without the extra preconditioner
// application of gmres.hpp:198-202 (FGMRES, new).
// ---------------------------------------------------------------------------
using ApplyFn = std::function<void(const DVec&, DVec&)>;

numa_gmres_result gmres(const ApplyFn& apply, const ApplyFn& precond, const DVec& b, DVec& x,
                        double rtol, size_t restart, size_t max_iters, bool zero_guess,
                        bool flexible, const DVec* r_init = nullptr,
                        std::vector<DVec>* saveV = nullptr,
                        std::vector<std::vector<double>>* saveH = nullptr, size_t* save_k = nullptr) {
    const size_t n = b.size();
    numa_gmres_result res{};
    if (save_k) *save_k = 0;
    const double bnorm = norm2(b);
    if (bnorm == 0.0) {
        x.assign(n, 0.0);
        res.converged = 1;
        return res;
    }
    if (restart == 0) {
        res.rel_res = 1.0;
        res.breakdown = 1;
        return res;
    }
    if (max_iters == 0) {
        res.rel_res = 1.0;
        return res;
    }
    if (zero_guess) x.assign(n, 0.0);

    DVec r(n), w(n), z(n), tmp(n);
    std::vector<DVec> V(restart + 1, DVec(n));
    std::vector<DVec> Z;
    if (flexible) Z.assign(restart, DVec(n));
    std::vector<std::vector<double>> H(restart + 1, std::vector<double>(restart, 0.0));
    std::vector<std::vector<double>> Hraw(restart + 1, std::vector<double>(restart, 0.0));
    std::vector<double> cs(restart), sn(restart), g(restart + 1), y(restart);
    bool use_given = r_init != nullptr;
    bool first = zero_guess;
    size_t k_last = 0;

    while ((size_t)res.iters < max_iters) {
        if (use_given) {
            r = *r_init;
            use_given = false;
            first = false;
        } else if (first) {
            r = b;
            first = false;
        } else {
            apply(x, w);
            ++res.matvecs;
            for (size_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
        }
        const double beta = norm2(r);
        res.rel_res = beta / bnorm;
        if (res.rel_res <= rtol) {
            res.converged = 1;
            k_last = 0;
            break;
        }
        for (auto& row : Hraw) std::fill(row.begin(), row.end(), 0.0);
        for (size_t i = 0; i < n; ++i) V[0][i] = r[i] / beta;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;

        size_t k = 0;
        k_last = 0;
        while (k < restart && (size_t)res.iters < max_iters) {
            const size_t j = k;
            DVec& zj = flexible ? Z[j] : z;
            precond(V[j], zj);
            ++res.precs;
            apply(zj, w);
            ++res.matvecs;
            for (size_t i = 0; i <= j; ++i) {
                H[i][j] = dot(w, V[i]);
                axpy(-H[i][j], V[i], w);
            }
            H[j + 1][j] = norm2(w);
            if (H[j + 1][j] > 0.0) {
                for (size_t i = 0; i < n; ++i) V[j + 1][i] = w[i] / H[j + 1][j];
            } else {
                std::fill(V[j + 1].begin(), V[j + 1].end(), 0.0);
            }
            for (size_t i = 0; i <= j + 1; ++i) Hraw[i][j] = H[i][j];
            for (size_t i = 0; i < j; ++i) {
                const double t = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
                H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
                H[i][j] = t;
            }
            const double denom = std::hypot(H[j][j], H[j + 1][j]);
            if (!(denom > 0.0) || !std::isfinite(denom)) {
                res.breakdown = 1;
                return res;
            }
            cs[j] = H[j][j] / denom;
            sn[j] = H[j + 1][j] / denom;
            H[j][j] = denom;
            H[j + 1][j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++k;
            ++res.iters;
            k_last = k;
            res.rel_res = std::abs(g[j + 1]) / bnorm;
            if (res.rel_res <= rtol) break;
        }
        for (size_t ii = k; ii-- > 0;) {
            const double pivot = H[ii][ii];
            if (!(std::abs(pivot) > 0.0) || !std::isfinite(pivot)) {
                res.breakdown = 1;
                return res;
            }
            double s = g[ii];
            for (size_t jj = ii + 1; jj < k; ++jj) s -= H[ii][jj] * y[jj];
            y[ii] = s / pivot;
        }
        if (flexible) {
            for (size_t i = 0; i < k; ++i) axpy(y[i], Z[i], x);
        } else {
            std::fill(tmp.begin(), tmp.end(), 0.0);
            for (size_t i = 0; i < k; ++i) axpy(y[i], V[i], tmp);
            precond(tmp, z);
            ++res.precs;
            axpy(1.0, z, x);
        }
        if (res.rel_res <= rtol) {
            res.converged = 1;
            break;
        }
    }
    if (saveV && k_last > 0) {  // gmres.hpp:210-220
        *save_k = k_last;
        *saveV = V;
        saveH->assign(k_last + 1, std::vector<double>(k_last, 0.0));
        for (size_t i = 0; i <= k_last; ++i)
            for (size_t j = 0; j < k_last; ++j) (*saveH)[i][j] = Hraw[i][j];
    }
    return res;
}

// gmres.hpp:229-323 (recycled second solve).
numa_gmres_result gmres_recycled(const ApplyFn& apply, const ApplyFn& precond, const DVec& b,
                                 DVec& x, double rtol, size_t restart, size_t max_iters,
                                 bool flexible, const std::vector<DVec>& V,
                                 const std::vector<std::vector<double>>& Hbar, size_t k) {
    const size_t n = b.size();
    x.assign(n, 0.0);
    if (k == 0 || V.size() < k + 1 || Hbar.size() < k + 1)
        return gmres(apply, precond, b, x, rtol, restart, max_iters, true, flexible);
    std::vector<double> rhs(k + 1, 0.0);
    for (size_t i = 0; i <= k; ++i) rhs[i] = dot(V[i], b);
    auto R = Hbar;
    std::vector<double> cs(k, 0.0), sn(k, 0.0);
    for (size_t j = 0; j < k; ++j) {
        for (size_t i = 0; i < j; ++i) {
            const double t = cs[i] * R[i][j] + sn[i] * R[i + 1][j];
            R[i + 1][j] = -sn[i] * R[i][j] + cs[i] * R[i + 1][j];
            R[i][j] = t;
        }
        const double denom = std::hypot(R[j][j], R[j + 1][j]);
        if (!(denom > 0.0) || !std::isfinite(denom))
            return gmres(apply, precond, b, x, rtol, restart, max_iters, true, flexible);
        cs[j] = R[j][j] / denom;
        sn[j] = R[j + 1][j] / denom;
        R[j][j] = denom;
        R[j + 1][j] = 0.0;
        const double t = cs[j] * rhs[j] + sn[j] * rhs[j + 1];
        rhs[j + 1] = -sn[j] * rhs[j] + cs[j] * rhs[j + 1];
        rhs[j] = t;
    }
    std::vector<double> y(k, 0.0);
    for (size_t ii = k; ii-- > 0;) {
        const double pivot = R[ii][ii];
        if (!(std::abs(pivot) > 0.0) || !std::isfinite(pivot)) {
            x.assign(n, 0.0);
            return gmres(apply, precond, b, x, rtol, restart, max_iters, true, flexible);
        }
        double s = rhs[ii];
        for (size_t jj = ii + 1; jj < k; ++jj) s -= R[ii][jj] * y[jj];
        y[ii] = s / pivot;
    }
    DVec residual = b, comb(n, 0.0), px(n, 0.0);
    for (size_t i = 0; i <= k; ++i) {
        double coef = 0.0;
        for (size_t j = 0; j < k; ++j) coef += Hbar[i][j] * y[j];
        axpy(-coef, V[i], residual);
    }
    for (size_t j = 0; j < k; ++j) axpy(y[j], V[j], comb);
    precond(comb, px);
    x = std::move(px);
    auto result = gmres(apply, precond, b, x, rtol, restart, max_iters, false, flexible, &residual);
    ++result.precs;
    return result;
}

// ---------------------------------------------------------------------------
// Nonlinear outer loops.  methods/krylov_newton.hpp:54-189.
// ---------------------------------------------------------------------------
double forcing_ew(const numa_krylov_params& p, size_t k, double rnorm, double rprev,
                  double eta_prev, double res0);

// The build's final-solve safeguard (no reference counterpart; off when
// tight_eta = 0, the reference's behaviour).  Below T = max(tight_res,
// rtol_nl) ||F(x0)|| the forcing term is at most tight_eta, and the solve is
// converged only once its last accepted step started below T.
double tight_threshold(const numa_krylov_params& p, double res0) {
    return std::max(p.tight_res, p.rtol_nl) * res0;
}

bool nl_converged(const numa_krylov_params& p, double rnorm, double res0, int64_t iters,
                  double last_start) {
    if (!(rnorm <= p.rtol_nl * res0)) return false;
    if (!(p.tight_eta > 0.0) || iters == 0) return true;
    return last_start <= tight_threshold(p, res0);
}

double forcing(const numa_krylov_params& p, size_t k, double rnorm, double rprev,
               double eta_prev, double res0) {
    const double eta = forcing_ew(p, k, rnorm, rprev, eta_prev, res0);
    if (p.tight_eta > 0.0 && rnorm <= tight_threshold(p, res0)) return std::min(eta, p.tight_eta);
    return eta;
}

double forcing_ew(const numa_krylov_params& p, size_t k, double rnorm, double rprev,
                  double eta_prev, double res0) {
    if (p.forcing == NUMA_FORCING_CONSTANT) return p.eta_init;
    // krylov_newton.hpp:54-65
    if (k == 0) return p.eta_init;
    double eta = p.gamma * std::pow(rnorm / rprev, p.alpha);
    const double safe = p.gamma * std::pow(eta_prev, p.alpha);
    if (safe > 0.1) eta = std::max(eta, safe);
    eta = std::min(eta, p.eta_max);
    eta = std::max(eta, 0.5 * p.rtol_nl * res0 / rnorm);
    return std::min(eta, p.eta_max);
}

// A problem as seen by the outer loops: F, J v at the current x, F''[v,v],
// and the preconditioner bound at x.
struct Problem {
    size_t n;
    bool jvp_needs_F = false;  // J v uses F at its base point (finite differences)
    std::function<void(const DVec&, DVec&)> eval;
    std::function<void(const DVec& x, const DVec& Fx, double xnorm, const DVec& v, DVec& out)> jvp;
    std::function<void(const DVec& x, const DVec& Fx, double xnorm, const DVec& v, DVec& out)> d2;
    std::function<ApplyFn(const DVec& x)> make_precond;
};

void outer_solve(const Problem& P, DVec& x, const numa_krylov_params& p, bool chebyshev,
                 bool flexible, numa_krylov_result& out) {
    const auto t0 = std::chrono::steady_clock::now();
    const size_t n = P.n;
    out = numa_krylov_result{};
    DVec Fx(n), Fn(n), z(n), u(n), w(n), xn(n);
    P.eval(x, Fx);
    ++out.n_F;
    double rnorm = norm2(Fx);
    out.res0 = rnorm;
    double rprev = rnorm, eta_prev = p.eta_init;
    for (size_t k = 0; k < (size_t)p.max_nl; ++k) {
        if (nl_converged(p, rnorm, out.res0, out.iters, rprev)) {
            out.converged = 1;
            break;
        }
        const double eta1 = forcing(p, k, rnorm, rprev, eta_prev, out.res0);
        const double eta2 = std::min(p.eta2_max, p.eta2_factor * eta1);
        const ApplyFn M = P.make_precond(x);
        const double xnorm = norm2(x);
        ApplyFn apply = [&](const DVec& v, DVec& o) { P.jvp(x, Fx, xnorm, v, o); };
        z.assign(n, 0.0);
        std::vector<DVec> sV;
        std::vector<std::vector<double>> sH;
        size_t sk = 0;
        const bool rec = chebyshev && p.recycle;
        const auto g1 = gmres(apply, M, Fx, z, eta1, (size_t)p.gmres_restart,
                              (size_t)p.gmres_max_iters, true, flexible, nullptr,
                              rec ? &sV : nullptr, rec ? &sH : nullptr, rec ? &sk : nullptr);
        out.n_JVP += g1.matvecs;
        out.n_lin_iters += g1.iters;
        out.n_prec += g1.precs;
        if (!g1.converged) ++out.n_lin_failures;
        if (g1.breakdown) break;

        bool ok = false;
        if (chebyshev) {  // krylov_newton.hpp:147-169
            P.d2(x, Fx, xnorm, z, w);
            ++out.n_D2;
            u.assign(n, 0.0);
            const auto g2 = rec ? gmres_recycled(apply, M, w, u, eta2, (size_t)p.gmres_restart,
                                                 (size_t)p.gmres_max_iters, flexible, sV, sH, sk)
                                : gmres(apply, M, w, u, eta2, (size_t)p.gmres_restart,
                                        (size_t)p.gmres_max_iters, true, flexible);
            out.n_JVP += g2.matvecs;
            out.n_lin_iters += g2.iters;
            out.n_prec += g2.precs;
            if (!g2.converged) ++out.n_lin_failures;
            if (!g2.breakdown) {
                for (size_t i = 0; i < n; ++i) xn[i] = x[i] - z[i] - 0.5 * u[i];
                P.eval(xn, Fn);
                ++out.n_F;
                if (norm2(Fn) < rnorm) ok = true;
            }
        }
        if (!ok) {  // Armijo backtracking, krylov_newton.hpp:96-104 / 170-178
            double lam = 1.0;
            while (lam >= p.min_lambda) {
                for (size_t i = 0; i < n; ++i) xn[i] = x[i] - lam * z[i];
                P.eval(xn, Fn);
                ++out.n_F;
                if (norm2(Fn) <= (1.0 - p.armijo * lam) * rnorm) {
                    ok = true;
                    break;
                }
                lam *= 0.5;
            }
        }
        if (!ok) break;
        x = xn;
        Fx = Fn;
        rprev = rnorm;
        rnorm = norm2(Fx);
        eta_prev = eta1;
        ++out.iters;
    }
    if (rnorm <= p.rtol_nl * out.res0) out.converged = 1;
    out.res = rnorm;
    out.t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Tensor-free Chebyshev of Eustaquio, Ribeiro and Dumett (2018) with
// alpha = 0, methods/tf_chebyshev_2018.hpp:23-132: d1 = -z from J z = F,
// curvature C d1 = [J(x + h d1) d1 - J(x) d1] / h, J d2 = -C d1 / 2,
// x+ = x + d1 + d2.
double tf2018_step_length(const DVec& x, const DVec& d1) {  // :23-29
    const double eps = std::numeric_limits<double>::epsilon();
    const double nd = norm2(d1);
    if (!(nd > 0.0)) return 2.0 * eps;
    const double raw = std::sqrt(eps) * norm2(x) / nd;
    return std::clamp(raw, 2.0 * eps, 0.5);
}

// :31-51.  J v at the shifted point x + h d1 takes that point's own F and
// norm (the finite-difference J v evaluates F at its base point).
void tf2018_curvature(const Problem& P, const DVec& x, const DVec& Fx, double xnorm,
                      const DVec& d1, DVec& out, double* h_used) {
    const size_t n = x.size();
    const double h = tf2018_step_length(x, d1);
    DVec xh = x;
    axpy(h, d1, xh);
    DVec Fh(n), j0(n), j1(n);
    if (P.jvp_needs_F) P.eval(xh, Fh);
    P.jvp(x, Fx, xnorm, d1, j0);
    P.jvp(xh, Fh, norm2(xh), d1, j1);
    out.resize(n);
    for (size_t i = 0; i < n; ++i) out[i] = (j1[i] - j0[i]) / h;
    if (h_used) *h_used = h;
}

void tf2018_solve(const Problem& P, DVec& x, const numa_krylov_params& p, bool flexible,
                  numa_krylov_result& out) {
    const auto t0 = std::chrono::steady_clock::now();
    const size_t n = P.n;
    out = numa_krylov_result{};
    DVec Fx(n), z(n), d1(n), d2(n), curv(n), rhs2(n);
    P.eval(x, Fx);
    ++out.n_F;
    double rnorm = norm2(Fx);
    out.res0 = rnorm;
    const double eps = std::numeric_limits<double>::epsilon();
    for (size_t k = 0; k < (size_t)p.max_nl; ++k) {
        if (rnorm <= p.rtol_nl * out.res0) {
            out.converged = 1;
            break;
        }
        const double eta1 = std::clamp(std::min(1e-8, rnorm * rnorm), 2.0 * eps, 1e-8);
        const double eta2 = std::clamp(std::min(1e-8, rnorm), 2.0 * eps, 1e-8);
        const ApplyFn M = P.make_precond(x);
        const double xnorm = norm2(x);
        ApplyFn apply = [&](const DVec& v, DVec& o) { P.jvp(x, Fx, xnorm, v, o); };
        z.assign(n, 0.0);
        const auto g1 = gmres(apply, M, Fx, z, eta1, (size_t)p.gmres_restart,
                              (size_t)p.gmres_max_iters, true, flexible);
        out.n_JVP += g1.matvecs;
        out.n_lin_iters += g1.iters;
        out.n_prec += g1.precs;
        if (!g1.converged) {
            ++out.n_lin_failures;
            break;
        }
        if (g1.breakdown) break;
        for (size_t i = 0; i < n; ++i) d1[i] = -z[i];
        tf2018_curvature(P, x, Fx, xnorm, d1, curv, nullptr);
        out.n_JVP += 2;
        for (size_t i = 0; i < n; ++i) rhs2[i] = -0.5 * curv[i];
        d2.assign(n, 0.0);
        const auto g2 = gmres(apply, M, rhs2, d2, eta2, (size_t)p.gmres_restart,
                              (size_t)p.gmres_max_iters, true, flexible);
        out.n_JVP += g2.matvecs;
        out.n_lin_iters += g2.iters;
        out.n_prec += g2.precs;
        if (!g2.converged) {
            ++out.n_lin_failures;
            break;
        }
        if (g2.breakdown) break;
        axpy(1.0, d1, x);
        axpy(1.0, d2, x);
        P.eval(x, Fx);
        ++out.n_F;
        rnorm = norm2(Fx);
        ++out.iters;
        if (!std::isfinite(rnorm)) break;
    }
    if (rnorm <= p.rtol_nl * out.res0) out.converged = 1;
    out.res = rnorm;
    out.t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------------------
// Smallest eigenpair by inverse iteration.  analysis/spectral.hpp:34-115.
// ---------------------------------------------------------------------------
numa_eigen_result inverse_smallest_eigenpair(const ApplyFn& apply, const ApplyFn& precond,
                                             size_t n, const numa_eigen_params& p, DVec* vec) {
    numa_eigen_result r{};
    r.eigenvalue = std::numeric_limits<double>::quiet_NaN();
    r.residual_rel = std::numeric_limits<double>::infinity();
    if (n == 0 || !(p.eigen_rtol > 0.0) || !(p.linear_rtol > 0.0) || p.max_outer <= 0 ||
        p.restart <= 0 || p.max_linear <= 0)
        return r;
    DVec q(n);
    for (size_t i = 0; i < n; ++i)
        q[i] = 1.0 + 0.1 * std::sin(0.6180339887498948 * static_cast<double>(i + 1));
    const double n0 = norm2(q);
    if (!(n0 > 0.0) || !std::isfinite(n0)) return r;
    for (double& v : q) v /= n0;
    double previous = std::numeric_limits<double>::quiet_NaN();
    DVec y(n), jq(n), res(n);
    for (int64_t outer = 0; outer < p.max_outer; ++outer) {
        y.assign(n, 0.0);
        const auto lin = gmres(apply, precond, q, y, p.linear_rtol, (size_t)p.restart,
                               (size_t)p.max_linear, true, false);
        r.linear_iterations += lin.iters;
        r.linear_matvecs += lin.matvecs;
        r.preconditioner_applications += lin.precs;
        if (!lin.converged || lin.breakdown) {
            ++r.linear_failures;
            return r;
        }
        const double yn = norm2(y);
        if (!(yn > std::numeric_limits<double>::min()) || !std::isfinite(yn)) {
            ++r.linear_failures;
            return r;
        }
        for (size_t i = 0; i < n; ++i) q[i] = y[i] / yn;
        apply(q, jq);
        const double theta = dot(q, jq);
        if (!std::isfinite(theta)) return r;
        res = jq;
        axpy(-theta, q, res);
        const double denom = std::max(1.0, std::abs(theta));
        const double rr = norm2(res) / denom;
        r.eigenvalue = theta;
        r.residual_rel = rr;
        r.outer_iterations = outer + 1;
        if (vec) *vec = q;
        const double change = std::isfinite(previous) ? std::abs(theta - previous) / denom
                                                      : std::numeric_limits<double>::infinity();
        if (rr <= p.eigen_rtol && (!std::isfinite(previous) || change <= 10.0 * p.eigen_rtol)) {
            r.converged = 1;
            return r;
        }
        previous = theta;
    }
    return r;
}

ApplyFn make_bratu_precond(const Bratu& f, const numa_solver_config& cfg,
                           std::shared_ptr<MG> mg, const DVec& x) {
    if (cfg.precond == NUMA_PRECOND_GMG) {
        return [mg](const DVec& r, DVec& z) {
            z.resize(r.size());
            mg->apply(r.data(), z.data());
        };
    }
    if (cfg.precond == NUMA_PRECOND_JACOBI) {  // bratu3d.hpp:104-120
        auto d = std::make_shared<DVec>(x.size());
        f.jacobi_diag(x.data(), d->data());
        return [d](const DVec& r, DVec& z) {
            z.resize(r.size());
            for (size_t i = 0; i < r.size(); ++i) z[i] = r[i] / (*d)[i];
        };
    }
    return [](const DVec& r, DVec& z) { z = r; };  // IdentityPrecond, gmres.hpp:37-39
}

// Fault injection of the failure-path tests (orc_set_inject): element 0 of
// the `at`-th F evaluation (what 0) or J v product (what 1) becomes NaN,
// once.  The same hook as the device solver's numa_hip_solver_inject_nan.
thread_local int g_inj_what = -1;
thread_local int64_t g_inj_at = 0, g_inj_count = 0;

void inject(int what, DVec& o) {
    if (g_inj_what != what || ++g_inj_count != g_inj_at) return;
    g_inj_what = -1;
    o[0] = std::numeric_limits<double>::quiet_NaN();
}

Problem bratu_problem(const Bratu& f, const numa_solver_config& cfg) {
    Problem P;
    P.n = (size_t)f.size();
    std::shared_ptr<MG> mg;
    if (cfg.precond == NUMA_PRECOND_GMG) mg = std::make_shared<MG>(f.n, cfg.mg);
    P.eval = [&f](const DVec& x, DVec& o) {
        o.resize(x.size());
        f.eval(x.data(), o.data());
        inject(NUMA_INJECT_F, o);
    };
    P.jvp_needs_F = cfg.jv_mode == NUMA_DERIV_FD;
    if (cfg.jv_mode == NUMA_DERIV_FD)
        P.jvp = [&f](const DVec& x, const DVec& Fx, double xn, const DVec& v, DVec& o) {
            o.resize(x.size());
            jvp_fd(f, x.data(), v.data(), o.data(), Fx.data(), xn);
            inject(NUMA_INJECT_JV, o);
        };
    else
        P.jvp = [&f](const DVec& x, const DVec&, double, const DVec& v, DVec& o) {
            o.resize(x.size());
            f.jvp(x.data(), v.data(), o.data());
            inject(NUMA_INJECT_JV, o);
        };
    if (cfg.d2_mode == NUMA_DERIV_FD)
        P.d2 = [&f](const DVec& x, const DVec& Fx, double xn, const DVec& v, DVec& o) {
            o.resize(x.size());
            d2_fd(f, x.data(), v.data(), o.data(), Fx.data(), xn);
        };
    else
        P.d2 = [&f](const DVec& x, const DVec&, double, const DVec& v, DVec& o) {
            o.resize(x.size());
            f.d2(x.data(), v.data(), o.data());
        };
    P.make_precond = [&f, cfg, mg](const DVec& x) { return make_bratu_precond(f, cfg, mg, x); };
    return P;
}

// ---------------------------------------------------------------------------
// Chandrasekhar H-equation.  problems/chandrasekhar_op.hpp:31-136 (rule 0 =
// Gauss-Legendre, form 0) and functions/gauss_legendre_dynamic.hpp:50-114.
// ---------------------------------------------------------------------------
void gauss_legendre_unit(size_t n, DVec& nodes, DVec& weights) {
    nodes.assign(n, 0.0);
    weights.assign(n, 0.0);
    const size_t half = (n + 1) / 2;
    const long double pi = std::acos(-1.0L);
    const long double eps = 32.0L * std::numeric_limits<long double>::epsilon();
    auto legendre = [n](long double x, long double& p, long double& dp) {
        long double p0 = 1.0L, p1 = x;
        if (n == 0) {
            p = 1.0L;
            dp = 0.0L;
            return;
        }
        for (size_t k = 2; k <= n; ++k) {
            const long double kd = (long double)k;
            const long double pk = ((2.0L * kd - 1.0L) * x * p1 - (kd - 1.0L) * p0) / kd;
            p0 = p1;
            p1 = pk;
        }
        p = p1;
        dp = (long double)n * (x * p1 - p0) / (x * x - 1.0L);
    };
    for (size_t i = 0; i < half; ++i) {
        long double z = std::cos(pi * ((long double)i + 0.75L) / ((long double)n + 0.5L));
        for (int it = 0; it < 100; ++it) {
            long double p, dp;
            legendre(z, p, dp);
            const long double delta = p / dp;
            z -= delta;
            if (std::abs(delta) <= eps * (1.0L + std::abs(z))) break;
        }
        long double p, dp;
        legendre(z, p, dp);
        const long double wpm = 2.0L / ((1.0L - z * z) * dp * dp);
        nodes[i] = (double)(0.5L * (1.0L - z));
        nodes[n - 1 - i] = (double)(0.5L * (1.0L + z));
        weights[i] = (double)(0.5L * wpm);
        weights[n - 1 - i] = (double)(0.5L * wpm);
    }
}

struct Chandra {
    size_t n;
    double c;
    int form;
    DVec mu, w;
    Chandra(size_t n_, double c_, int rule, int form_) : n(n_), c(c_), form(form_) {
        if (rule == 0) {
            gauss_legendre_unit(n, mu, w);
        } else {  // midpoint: mu_i = (i - 1/2)/N for i = 1..N, w = 1/N
            mu.resize(n);
            w.assign(n, 1.0 / double(n));
            for (size_t i = 0; i < n; ++i) mu[i] = (double(i) + 0.5) / double(n);
        }
    }
    // chandrasekhar_op.hpp:109-127 (same product order)
    void apply_A(const DVec& v, DVec& out) const {
        out.assign(n, 0.0);
        const long nl = (long)n;
#pragma omp parallel for schedule(static) if (nl > 256)
        for (long ii = 0; ii < nl; ++ii) {
            const double mi = mu[(size_t)ii];
            double s = 0.0;
            for (size_t j = 0; j < n; ++j) s += 0.5 * c * w[j] * mi / (mi + mu[j]) * v[j];
            out[(size_t)ii] = s;
        }
    }
    void eval(const DVec& x, DVec& out) const {
        DVec Ax;
        apply_A(x, Ax);
        out.resize(n);
        for (size_t i = 0; i < n; ++i)
            out[i] = form == 0 ? 1.0 + x[i] * Ax[i] - x[i]  // chandrasekhar_op.hpp:56-66
                               : x[i] - 1.0 / (1.0 - Ax[i]);
    }
    void jvp(const DVec& x, const DVec& v, DVec& out) const {
        DVec Ax, Av;
        apply_A(x, Ax);
        apply_A(v, Av);
        out.resize(n);
        for (size_t i = 0; i < n; ++i) {
            if (form == 0) {
                out[i] = v[i] * Ax[i] + x[i] * Av[i] - v[i];  // :68-79
            } else {
                const double q = 1.0 - Ax[i];
                out[i] = v[i] - Av[i] / (q * q);
            }
        }
    }
    void d2(const DVec& x, const DVec& v, DVec& out) const {
        DVec Av, Ax;
        apply_A(v, Av);
        if (form != 0) apply_A(x, Ax);
        out.resize(n);
        for (size_t i = 0; i < n; ++i) {
            if (form == 0) {
                out[i] = 2.0 * v[i] * Av[i];  // :81-91
            } else {
                const double q = 1.0 - Ax[i];
                out[i] = -2.0 * Av[i] * Av[i] / (q * q * q);
            }
        }
    }
};

numa_mg_config default_mg() {
    numa_mg_config m{};
    m.smoother = NUMA_SMOOTH_RBGS;
    m.pre = 2;
    m.post = 2;
    m.degree = 3;
    m.coarse_iters = 30;
    m.min_coarse_n = 1;
    m.omega = 6.0 / 7.0;
    m.cheb_ratio = 4.0;
    return m;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void orc_set_inject(int64_t what, int64_t at) {
    g_inj_what = at > 0 ? (int)what : -1;
    g_inj_at = at;
    g_inj_count = 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_bratu_eval(int64_t n, double lambda, const double* u, double* out) { Bratu(n, lambda).eval(u, out); }
void orc_bratu_jvp(int64_t n, double lambda, const double* u, const double* v, double* out) {
    Bratu(n, lambda).jvp(u, v, out);
}
void orc_bratu_d2(int64_t n, double lambda, const double* u, const double* v, double* out) {
    Bratu(n, lambda).d2(u, v, out);
}
void orc_bratu_jacobi_diag(int64_t n, double lambda, const double* u, double* d) {
    Bratu(n, lambda).jacobi_diag(u, d);
}
void orc_bratu_jvp_fd(int64_t n, double lambda, const double* x, const double* v, double* out) {
    Bratu f(n, lambda);
    jvp_fd(f, x, v, out);
}
void orc_bratu_d2_fd(int64_t n, double lambda, const double* x, const double* v, double* out) {
    Bratu f(n, lambda);
    d2_fd(f, x, v, out);
}
void orc_poisson_apply(int64_t n, const double* v, double* out) { Bratu(n, 0.0).lap(v, out); }

double orc_dot(int64_t n, const double* a, const double* b) { return dot(a, b, (long)n); }

int64_t orc_mg_levels(int64_t n, const numa_mg_config* cfg, int64_t* sizes) {
    MG mg((long)n, cfg ? *cfg : default_mg());
    if (sizes)
        for (size_t l = 0; l < mg.levels.size(); ++l) sizes[l] = mg.levels[l].n;
    return (int64_t)mg.levels.size();
}

int orc_mg_apply(int64_t n, const numa_mg_config* cfg, const double* r, double* z) {
    if (n <= 0) return NUMA_ERR_ARG;
    MG mg((long)n, cfg ? *cfg : default_mg());
    mg.apply(r, z);
    return NUMA_OK;
}

int orc_bratu_gmres(int64_t n, double lambda, const numa_solver_config* cfg, const double* x0,
                    const double* b, double* x, double rtol, int64_t restart, int64_t max_iters,
                    int64_t zero_guess, numa_gmres_result* out) {
    if (n <= 0 || !cfg || !out) return NUMA_ERR_ARG;
    Bratu f((long)n, lambda);
    const size_t N = (size_t)f.size();
    DVec X0(x0, x0 + N), B(b, b + N), X(x, x + N), F0(N);
    f.eval(X0.data(), F0.data());
    const double xnorm = norm2(X0);
    Problem P = bratu_problem(f, *cfg);
    ApplyFn M = P.make_precond(X0);
    ApplyFn apply = [&](const DVec& v, DVec& o) { P.jvp(X0, F0, xnorm, v, o); };
    *out = gmres(apply, M, B, X, rtol, (size_t)std::max<int64_t>(restart, 0),
                 (size_t)std::max<int64_t>(max_iters, 0), zero_guess != 0, cfg->flexible != 0);
    std::copy(X.begin(), X.end(), x);
    return NUMA_OK;
}

int orc_poisson_gmres(int64_t n, const numa_mg_config* mgc, int64_t use_mg, const double* b,
                      double* x, double rtol, int64_t restart, int64_t max_iters,
                      numa_gmres_result* out) {
    if (n <= 0 || !out) return NUMA_ERR_ARG;
    Bratu f((long)n, 0.0);
    const size_t N = (size_t)f.size();
    auto mg = std::make_shared<MG>((long)n, mgc ? *mgc : default_mg());
    ApplyFn apply = [&](const DVec& v, DVec& o) {
        o.resize(N);
        f.lap(v.data(), o.data());
    };
    ApplyFn M = use_mg ? ApplyFn([mg](const DVec& r, DVec& z) {
        z.resize(r.size());
        mg->apply(r.data(), z.data());
    })
                       : ApplyFn([](const DVec& r, DVec& z) { z = r; });
    DVec B(b, b + N), X(N, 0.0);
    *out = gmres(apply, M, B, X, rtol, (size_t)restart, (size_t)max_iters, true, false);
    std::copy(X.begin(), X.end(), x);
    return NUMA_OK;
}

int orc_bratu_solve(int64_t n, double lambda, const numa_solver_config* cfg,
                    const numa_krylov_params* p, double* x, numa_krylov_result* out) {
    if (n <= 0 || !cfg || !p || !out) return NUMA_ERR_ARG;
    Bratu f((long)n, lambda);
    Problem P = bratu_problem(f, *cfg);
    DVec X(x, x + f.size());
    if (cfg->method == NUMA_METHOD_TF2018)
        tf2018_solve(P, X, *p, cfg->flexible != 0, *out);
    else
        outer_solve(P, X, *p, cfg->method == NUMA_METHOD_CHEBYSHEV, cfg->flexible != 0, *out);
    std::copy(X.begin(), X.end(), x);
    return NUMA_OK;
}

int orc_bratu_eigen_smallest(int64_t n, double lambda, const numa_solver_config* cfg,
                             const double* x, const numa_eigen_params* p, numa_eigen_result* out,
                             double* vec) {
    if (n <= 0 || !cfg || !x || !p || !out) return NUMA_ERR_ARG;
    Bratu f((long)n, lambda);
    Problem P = bratu_problem(f, *cfg);
    const DVec X(x, x + f.size());
    DVec Fx(f.size());
    if (P.jvp_needs_F) P.eval(X, Fx);
    const double xn = norm2(X);
    ApplyFn apply = [&](const DVec& v, DVec& o) { P.jvp(X, Fx, xn, v, o); };
    DVec q;
    *out = inverse_smallest_eigenpair(apply, P.make_precond(X), f.size(), *p, &q);
    if (vec && q.size() == (size_t)f.size()) std::copy(q.begin(), q.end(), vec);
    return NUMA_OK;
}

// Natural-parameter continuation with the spectral stop, apps/calibrate_b6.cpp:
// 139-171 (fold proxy) and 173-298 (loop).
int orc_bratu_continuation(int64_t n, const numa_solver_config* cfg,
                           const numa_continuation_params* cp, const numa_krylov_params* nk,
                           numa_continuation_point* points, int64_t max_points, int64_t* count,
                           int64_t* status, double* fold_proxy, double* x_out) {
    if (n <= 0 || !cfg || !cp || !nk || !count || !status || !fold_proxy) return NUMA_ERR_ARG;
    const size_t N = (size_t)(n * n * n);
    DVec previous(N, 0.0);
    double lam = cp->lambda_start, step = cp->step;
    std::vector<numa_continuation_point> acc;
    *status = 0;
    // one multigrid hierarchy for every lambda (PoissonMGProvider, :179)
    numa_solver_config c = *cfg;
    std::shared_ptr<MG> mg;
    if (c.precond == NUMA_PRECOND_GMG) mg = std::make_shared<MG>((long)n, c.mg);
    while (lam <= cp->lambda_stop + 1e-14) {
        Bratu f((long)n, lam);
        Problem P = bratu_problem(f, c);
        if (mg) P.make_precond = [&f, c, mg](const DVec& x) { return make_bratu_precond(f, c, mg, x); };
        DVec x = previous;
        numa_krylov_result sol{};
        outer_solve(P, x, *nk, false, c.flexible != 0, sol);
        if (!sol.converged) {
            step *= 0.5;
            if (step < cp->min_step) {
                *status = 3;
                break;
            }
            lam = acc.empty() ? cp->lambda_start : acc.back().lambda + step;
            continue;
        }
        DVec Fx(N);
        if (P.jvp_needs_F) P.eval(x, Fx);
        const double xn = norm2(x);
        ApplyFn apply = [&](const DVec& v, DVec& o) { P.jvp(x, Fx, xn, v, o); };
        const numa_eigen_params ep{cp->eigen_rtol, cp->linear_rtol, 30, 100, 2000};
        const numa_eigen_result eig =
            inverse_smallest_eigenpair(apply, P.make_precond(x), N, ep, nullptr);
        if (!eig.converged || !std::isfinite(eig.eigenvalue) || eig.eigenvalue <= 0.0) {
            *status = 4;
            break;
        }
        acc.push_back({lam, eig.eigenvalue, eig.residual_rel, sol.iters, sol.n_F,
                       sol.n_JVP + eig.linear_matvecs, sol.n_prec + eig.preconditioner_applications});
        previous = x;
        if (eig.eigenvalue <= cp->spectral_stop) break;
        if (sol.iters <= 5 && step < cp->step) step = std::min(cp->step, 1.5 * step);
        lam += step;
    }
    *count = (int64_t)acc.size();
    for (size_t i = 0; i < acc.size() && (int64_t)i < max_points; ++i) points[i] = acc[i];
    // fold proxy: linear fit of gap^2 over the last four points (:139-171)
    *fold_proxy = std::numeric_limits<double>::quiet_NaN();
    if (acc.size() >= 4) {
        const size_t first = acc.size() - 4;
        double sx = 0.0, sy = 0.0, sxx = 0.0, sxy = 0.0;
        for (size_t i = first; i < acc.size(); ++i) {
            const double xv = acc[i].lambda, yv = acc[i].spectral_gap * acc[i].spectral_gap;
            sx += xv;
            sy += yv;
            sxx += xv * xv;
            sxy += xv * yv;
        }
        const double cnt = 4.0, den = cnt * sxx - sx * sx;
        if (std::abs(den) >= 1e-14) {
            const double slope = (cnt * sxy - sx * sy) / den;
            const double icpt = (sy - slope * sx) / cnt;
            const double root = -icpt / slope;
            if (slope < 0.0 && std::isfinite(root) && root > acc.back().lambda) *fold_proxy = root;
        }
    }
    if (x_out) std::copy(previous.begin(), previous.end(), x_out);
    return NUMA_OK;
}

int orc_tf2018_curvature(int64_t n, double lambda, int64_t jv_mode, const double* x,
                         const double* d, double* out, double* h_used) {
    if (n <= 0 || !x || !d || !out) return NUMA_ERR_ARG;
    Bratu f((long)n, lambda);
    numa_solver_config cfg{};
    cfg.jv_mode = jv_mode;
    Problem P = bratu_problem(f, cfg);
    const DVec X(x, x + f.size()), D(d, d + f.size());
    DVec Fx(f.size()), O;
    if (P.jvp_needs_F) P.eval(X, Fx);
    tf2018_curvature(P, X, Fx, norm2(X), D, O, h_used);
    std::copy(O.begin(), O.end(), out);
    return NUMA_OK;
}

int orc_gauss_legendre_unit(int64_t n, double* nodes, double* weights) {
    if (n <= 0) return NUMA_ERR_ARG;
    DVec a, b;
    gauss_legendre_unit((size_t)n, a, b);
    std::copy(a.begin(), a.end(), nodes);
    std::copy(b.begin(), b.end(), weights);
    return NUMA_OK;
}

int orc_chandra_eval(int64_t n, double c, int64_t rule, int64_t form, const double* x, double* out) {
    if (n <= 0) return NUMA_ERR_ARG;
    Chandra op((size_t)n, c, (int)rule, (int)form);
    DVec X(x, x + n), O;
    op.eval(X, O);
    std::copy(O.begin(), O.end(), out);
    return NUMA_OK;
}
int orc_chandra_jvp(int64_t n, double c, int64_t rule, int64_t form, const double* x,
                    const double* v, double* out) {
    if (n <= 0) return NUMA_ERR_ARG;
    Chandra op((size_t)n, c, (int)rule, (int)form);
    DVec X(x, x + n), Vv(v, v + n), O;
    op.jvp(X, Vv, O);
    std::copy(O.begin(), O.end(), out);
    return NUMA_OK;
}
int orc_chandra_d2(int64_t n, double c, int64_t rule, int64_t form, const double* x,
                   const double* v, double* out) {
    if (n <= 0) return NUMA_ERR_ARG;
    Chandra op((size_t)n, c, (int)rule, (int)form);
    DVec X(x, x + n), Vv(v, v + n), O;
    op.d2(X, Vv, O);
    std::copy(O.begin(), O.end(), out);
    return NUMA_OK;
}

int orc_chandra_solve(int64_t n, double c, int64_t rule, int64_t form, const numa_solver_config* cfg,
                      const numa_krylov_params* p, double* x, numa_krylov_result* out) {
    if (n <= 0 || !cfg || !p || !out) return NUMA_ERR_ARG;
    auto op = std::make_shared<Chandra>((size_t)n, c, (int)rule, (int)form);
    Problem P;
    P.n = (size_t)n;
    P.eval = [op](const DVec& X, DVec& o) { op->eval(X, o); };
    P.jvp = [op](const DVec& X, const DVec&, double, const DVec& v, DVec& o) { op->jvp(X, v, o); };
    P.d2 = [op](const DVec& X, const DVec&, double, const DVec& v, DVec& o) { op->d2(X, v, o); };
    P.make_precond = [](const DVec&) { return ApplyFn([](const DVec& r, DVec& z) { z = r; }); };
    DVec X(x, x + n);
    outer_solve(P, X, *p, cfg->method == NUMA_METHOD_CHEBYSHEV, cfg->flexible != 0, *out);
    std::copy(X.begin(), X.end(), x);
    return NUMA_OK;
}

}  // extern "C"
