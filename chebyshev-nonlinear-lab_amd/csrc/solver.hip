// solver.hip - device-resident (F)GMRES and Newton / inexact-Chebyshev outer
// loops on Bratu3D (methods/krylov_newton.hpp:54-189, linalg/gmres.hpp:61-323),
// plus the C entry points of the multigrid and solver handles.
//
// Every N-sized vector lives in HBM.  Per Arnoldi step the host launches
//   preconditioner -> ||z|| -> J z (FD, fused first MGS dot) -> j+1 fused MGS
// passes, then reads the j+2 new Hessenberg entries (one small copy) and
// applies the Givens rotations in FP64, exactly in the reference's order.
// Basis vectors are stored unnormalised with device-side scales s_i
// (V_i = s_i * Vraw_i), which removes the normalisation pass.
//
// Precision: T = double (FP64) or float (mixed).  In mixed mode the Krylov
// basis, Z vectors, J v output and the V-cycle are FP32; F(x), the Newton /
// Chebyshev update, the linear iterate, the Givens / least-squares solve and
// all dot products (FP64 accumulation) stay FP64.  The mixed FD J v evaluates
// the difference quotient in the rearranged form h2inv A v - lambda e^x
// (e^{h v} - 1) / h in FP64 arithmetic from v and a FP32 copy of lambda e^x.
#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <tuple>

#include "blas.hpp"
#include "mg.hpp"
#include "solver_base.hpp"
#include "stencil.hpp"

struct numa_hip_mg {
    numa_hip::Ctx* ctx;
    int64_t precision;
    std::unique_ptr<numa_hip::MGBase> impl;
};

namespace numa_hip {

template <bool VEC>
__global__ __launch_bounds__(kBlock) void k_xnorm_exp(int64_t N, const double* x, double lambda,
                                                      const double* lam_ptr, float* E, Reduce red) {
    xnorm_exp_body<VEC>(hw_grid(), N, x, lambda, lam_ptr, E, red);
}

template <class T>
class BratuSolver : public SolverBase {
public:
    BratuSolver(Ctx& c, int64_t n, double lambda, const numa_solver_config& cfg, int64_t restart)
        : c_(c), n_(n), N_(n * n * n), lambda_(lambda), h2inv_((n + 1.0) * (n + 1.0)), cfg_(cfg),
          m_(restart) {
        require(n > 0, "Bratu3D needs n > 0");
        require(n <= 2048, "grid too large", NUMA_ERR_OVERFLOW);
        require(restart > 0 && restart <= 256, "gmres restart must be in [1, 256]");
        require(cfg.precond >= NUMA_PRECOND_NONE && cfg.precond <= NUMA_PRECOND_LEARNED,
                "unknown preconditioner");
        require(cfg.precond != NUMA_PRECOND_LEARNED || cfg.flexible,
                "the learned preconditioner is nonlinear and needs FGMRES (flexible = 1)");
        require(cfg.method >= NUMA_METHOD_NEWTON && cfg.method <= NUMA_METHOD_TF2018,
                "unknown method");
        if (cfg.precond == NUMA_PRECOND_GMG || cfg.precond == NUMA_PRECOND_LEARNED)
            mg_.reset(make_mg(c, n, cfg.mg, cfg.precision));
        x2_ = allocd();
        Fx_ = allocd();
        Fn_ = allocd();
        z_ = allocd();
        u_ = allocd();
        w_ = allocd();
        r_ = allocd();
        if (cfg.precond == NUMA_PRECOND_JACOBI) diag_ = allocd();
        if (sizeof(T) == 4 && cfg.jv_mode == NUMA_DERIV_FD) {  // lambda e^x of EpJvFDe
            void* pe = nullptr;
            NUMA_HIP_CHECK(hipMalloc(&pe, (size_t)N_ * sizeof(float)));
            allocs_.push_back(pe);
            E_ = static_cast<float*>(pe);
        }
        if (cfg.method == NUMA_METHOD_TF2018) {
            tfa_ = allocd();
            tfb_ = allocd();
        }
        V_.resize((size_t)m_ + 1);
        for (auto& v : V_) v = alloct();
        if (cfg.flexible && cfg.precond != NUMA_PRECOND_NONE) {
            Z_.resize((size_t)m_);
            for (auto& v : Z_) v = alloct();
        }
        zb_ = alloct();
        comb_ = alloct();
        // device scalars: H ((m+1) x m, column major), s (m+1), rhs (m+1), misc
        nsc_ = (m_ + 1) * m_ + 2 * (m_ + 1) + kMisc;
        void* p = nullptr;
        NUMA_HIP_CHECK(hipMalloc(&p, (size_t)nsc_ * sizeof(double)));
        allocs_.push_back(p);
        sd_ = static_cast<double*>(p);
        NUMA_HIP_CHECK(hipHostMalloc(&p, (size_t)nsc_ * sizeof(double), hipHostMallocDefault));
        hs_ = static_cast<double*>(p);
        launch_fill<double>(c_, 1, lambda_, MISC(S_LAMBDA));
    }
    ~BratuSolver() override {
        drop_graphs();
        for (auto& t : tpend_) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        for (hipEvent_t e : evpool_) (void)hipEventDestroy(e);
        for (void* p : allocs_) (void)hipFree(p);
        if (hs_) (void)hipHostFree(hs_);
    }
    int64_t size() const override { return N_; }
    int64_t grid_n() const override { return n_; }
    void set_lambda(double lambda) override {
        require(std::isfinite(lambda), "lambda must be finite");
        lambda_ = lambda;
        launch_fill<double>(c_, 1, lambda_, MISC(S_LAMBDA));
    }
    void set_graphs(bool on) override {
        if (on && !xa_) xa_ = allocd();
        if (!on) drop_graphs();
        graphs_ = on;
    }
    int64_t graphs_cached() const override { return (int64_t)gcache_.size(); }
    void set_timing(bool on) override {
        c_.sync();
        collect_timers();
        timing_ = on;
    }
    // analysis/spectral.hpp:34-115 on the device: GMRES on J(x) with the
    // configured preconditioner, FP64 J v, dots and norms on device scalars.
    numa_eigen_result eigen_smallest(const double* x, const numa_eigen_params& p,
                                     double* vec) override {
        require(sizeof(T) == 8, "the spectral estimator needs an FP64 solver");
        numa_eigen_result r{};
        r.eigenvalue = std::numeric_limits<double>::quiet_NaN();
        r.residual_rel = std::numeric_limits<double>::infinity();
        if (!(p.eigen_rtol > 0.0) || !(p.linear_rtol > 0.0) || p.max_outer <= 0 ||
            p.restart <= 0 || p.max_linear <= 0)
            return r;
        require(p.restart <= m_, "eigen restart exceeds the solver's basis capacity");
        point_setup(x);
        norm(x, MISC(S_XNORM));
        if (cfg_.jv_mode == NUMA_DERIV_FD) eval_F(x, nullptr, 0.0, nullptr, 0.0, Fx_, nullptr);
        double* q = w_;
        double* y = z_;
        double* jq = u_;
        double* res = r_;
        launch_eig_start(c_, N_, q);
        norm(q, MISC(S_TMP));
        const double n0 = read1(MISC(S_TMP));
        if (!(n0 > 0.0) || !std::isfinite(n0)) return r;
        launch_div_to(c_, N_, q, n0, q);
        double previous = std::numeric_limits<double>::quiet_NaN();
        for (int64_t outer = 0; outer < p.max_outer; ++outer) {
            const auto lin = gmres(x, Fx_, q, y, p.linear_rtol, p.restart, p.max_linear, true,
                                   nullptr, false);
            r.linear_iterations += lin.iters;
            r.linear_matvecs += lin.matvecs;
            r.preconditioner_applications += lin.precs;
            if (!lin.converged || lin.breakdown) {
                ++r.linear_failures;
                break;
            }
            norm(y, MISC(S_TMP));
            const double yn = read1(MISC(S_TMP));
            if (!(yn > std::numeric_limits<double>::min()) || !std::isfinite(yn)) {
                ++r.linear_failures;
                break;
            }
            launch_div_to(c_, N_, y, yn, q);
            jv64(x, Fx_, MISC(S_XNORM), q, jq);
            launch_dot<double, double>(c_, N_, q, jq,
                                       Epilogue{EPI_STORE, 1.0, nullptr, MISC(S_TMP), nullptr});
            const double theta = read1(MISC(S_TMP));
            if (!std::isfinite(theta)) break;
            launch_axpy_to(c_, N_, jq, -theta, q, res);  // J q - theta q
            norm(res, MISC(S_TMP));
            const double denom = std::max(1.0, std::abs(theta));
            const double rr = read1(MISC(S_TMP)) / denom;
            r.eigenvalue = theta;
            r.residual_rel = rr;
            r.outer_iterations = outer + 1;
            if (vec) launch_convert<double, double>(c_, N_, q, vec);
            const double change = std::isfinite(previous)
                                      ? std::abs(theta - previous) / denom
                                      : std::numeric_limits<double>::infinity();
            if (rr <= p.eigen_rtol &&
                (!std::isfinite(previous) || change <= 10.0 * p.eigen_rtol)) {
                r.converged = 1;
                break;
            }
            previous = theta;
        }
        c_.sync();
        collect_timers();
        return r;
    }

    double tf2018_curvature(const double* x, const double* d, double* out) override {
        require(n_ > 0, "empty problem");
        const double h = tf_curvature(x, nullptr, d, out, 1.0, x2_, Fn_);
        c_.sync();
        return h;
    }
    numa_timers timers(bool reset) override {
        c_.sync();
        collect_timers();
        const numa_timers t = tacc_;
        if (reset) tacc_ = numa_timers{};
        return t;
    }
    void set_capture(CaptureFn fn) override { capture_ = std::move(fn); }
    void set_learned(std::shared_ptr<NetApply> net, double ratio) override {
        require(cfg_.precond == NUMA_PRECOND_LEARNED,
                "solver was not created with NUMA_PRECOND_LEARNED");
        require(ratio > 0.0 && ratio <= 1.0, "fallback ratio must be in (0, 1]");
        net_ = std::move(net);
        ratio_ = ratio;
    }
    LearnedStats learned_stats() const override { return lstats_; }
    void inject_nan(int64_t what, int64_t call) override {
        require(what == NUMA_INJECT_F || what == NUMA_INJECT_JV, "unknown injection target");
        require(call >= 0, "negative call index");
        inj_what_ = call > 0 ? what : -1;
        inj_at_ = call;
        inj_count_ = 0;
    }
    const MGBase* multigrid() const override { return mg_.get(); }
    MGBase* multigrid() override { return mg_.get(); }

    numa_gmres_result gmres_entry(const double* x0, const double* b, double* z, double rtol,
                                  int64_t restart, int64_t max_iters, bool zero_guess,
                                  const double* r_init, bool save) override {
        point_setup(x0);
        eval_F(x0, nullptr, 0.0, nullptr, 0.0, Fx_, nullptr);
        auto r = gmres(x0, Fx_, b, z, rtol, restart, max_iters, zero_guess, r_init, save);
        c_.sync();
        collect_timers();
        return r;
    }

    int64_t saved_space(double* hbar, double* const* V) override {
        const int64_t k = saved_k_;
        if (hbar)
            for (int64_t i = 0; i <= k && k > 0; ++i)
                for (int64_t j = 0; j < k; ++j) hbar[i * k + j] = saved_H_[(size_t)i][(size_t)j];
        if (V && k > 0) {
            for (int64_t i = 0; i <= k; ++i) {  // V_i = s_i Vraw_i
                require(V[i] != nullptr, "null basis vector");
                launch_convert<T, double>(c_, N_, V_[(size_t)i], V[i]);
                launch_scale_to(c_, N_, saved_s_[(size_t)i], V[i], V[i]);
            }
            c_.sync();
        }
        return k;
    }

    numa_gmres_result gmres_recycled_entry(const double* x0, const double* b, double* z,
                                           double rtol, int64_t restart, int64_t max_iters,
                                           int64_t k, const double* hbar,
                                           const double* const* V, bool save) override {
        if (V) {  // load the given relation as the saved one (unit scales)
            require(k >= 0 && k <= m_, "saved space larger than the solver's basis");
            require(k == 0 || hbar != nullptr, "null Hbar");
            saved_k_ = k;
            saved_H_.assign((size_t)k + 1, std::vector<double>((size_t)k, 0.0));
            saved_s_.assign((size_t)k + 1, 1.0);
            for (int64_t i = 0; i <= k && k > 0; ++i) {
                require(V[i] != nullptr, "null basis vector");
                for (int64_t j = 0; j < k; ++j) saved_H_[(size_t)i][(size_t)j] = hbar[i * k + j];
                launch_convert<double, T>(c_, N_, V[i], V_[(size_t)i]);
            }
        }
        point_setup(x0);
        eval_F(x0, nullptr, 0.0, nullptr, 0.0, Fx_, nullptr);
        auto r = gmres_recycled(x0, Fx_, b, z, rtol, restart, max_iters, save);
        c_.sync();
        collect_timers();
        return r;
    }

    void solve(const numa_krylov_params& p, double* xdev, numa_krylov_result& out) override;

    // Average device time of one primitive call (HIP events on the solver's
    // stream) and its algorithmic HBM bytes; the counterpart of the
    // reference's frozen per-call costs c_F, c_JVP, c_D2, c_prec.
    double primitive_cost(int64_t prim, int64_t reps, double* bytes) override {
        require(reps > 0, "reps must be positive");
        const double N = (double)N_, st = sizeof(T);
        // well-defined inputs: x = 0.1, F = F(x), Krylov vectors = 1
        launch_fill<double>(c_, N_, 0.1, x2_);
        eval_F(x2_, nullptr, 0.0, nullptr, 0.0, Fx_, nullptr);
        launch_fill<double>(c_, N_, 0.01, z_);
        for (int i = 0; i < 3 && i <= m_; ++i) launch_fill<T>(c_, N_, T(1), V_[(size_t)i]);
        launch_fill<T>(c_, N_, T(1), zb_);
        point_setup(x2_);
        norm(zb_, MISC(S_VNORM));
        launch_fill<double>(c_, 1, 1.0, S(0));
        launch_fill<double>(c_, 1, 1.0, H(0, 0));
        auto run = [&]() -> double {
            switch (prim) {
                case NUMA_PRIM_F:
                    eval_F(x2_, nullptr, 0.0, nullptr, 0.0, Fn_, nullptr);
                    return 16.0 * N;
                case NUMA_PRIM_JVP: {  // stencil kernel only (||v|| precomputed)
                    if (cfg_.jv_mode == NUMA_DERIV_FD && E_) {
                        const FDStep stp{std::sqrt(std::numeric_limits<double>::epsilon()),
                                         MISC(S_XNORM), MISC(S_VNORM)};
                        launch_stencil<true>(c_, n_, LdVec<T>{zb_},
                                             EpJvFDe<T, true, false>{h2inv_, stp, S(0), E_, V_[1],
                                                                     V_[0], nullptr, 0.0, 1.0},
                                             Epilogue{EPI_STORE, 1.0, S(0), H(0, 0), nullptr});
                        return N * (4.0 + 3.0 * st);  // lambda e^x, v, out, V0
                    }
                    if (cfg_.jv_mode == NUMA_DERIV_FD) {
                        const FDStep stp{std::sqrt(std::numeric_limits<double>::epsilon()),
                                         MISC(S_XNORM), MISC(S_VNORM)};
                        launch_stencil<true>(c_, n_, LdXpHV<T>{x2_, zb_, stp, 0.0},
                                             EpJvFD<T, true, false>{h2inv_, lambda_, stp, S(0), Fx_,
                                                                    V_[1], V_[0], nullptr, 0.0, 1.0, lam_dev()},
                                             Epilogue{EPI_STORE, 1.0, S(0), H(0, 0), nullptr});
                        return N * (8.0 + 8.0 + 3.0 * st);  // x, F0, v, out, V0
                    }
                    launch_stencil<true>(c_, n_, LdVec<T>{zb_},
                                         EpJvAn<T, true, false>{h2inv_, lambda_, S(0), x2_, V_[1],
                                                                V_[0], nullptr, 1.0, lam_dev()},
                                         Epilogue{EPI_STORE, 1.0, S(0), H(0, 0), nullptr});
                    return N * (8.0 + 3.0 * st);
                }
                case NUMA_PRIM_D2:
                    d2(x2_, Fx_, z_, w_);
                    return N * 8.0 * (cfg_.d2_mode == NUMA_DERIV_FD ? 4.0 : 3.0) +
                           (cfg_.d2_mode == NUMA_DERIV_FD ? 8.0 * N : 0.0);
                case NUMA_PRIM_PREC:
                    precond(V_[0], zb_);
                    return mg_ ? mg_->vcycle_bytes() : 3.0 * N * st;
                case NUMA_PRIM_SMOOTH:
                    require(mg_ != nullptr, "no multigrid in this solver");
                    return mg_->smooth_step_for_timing();
                case NUMA_PRIM_COPY:  // the streaming copy that calibrates the HBM peak
                    launch_convert<double, double>(c_, N_, x2_, z_);
                    return 16.0 * N;
                case NUMA_PRIM_MGS:
                    launch_mgs<T>(c_, N_, V_[2], V_[0], H(0, 0), S(0), V_[1],
                                  Epilogue{EPI_STORE, 1.0, nullptr, MISC(S_TMP), nullptr});
                    return 4.0 * N * st;
                default:
                    throw Error(NUMA_ERR_ARG, "unknown primitive");
            }
        };
        require(prim != NUMA_PRIM_MGS || m_ >= 2, "MGS timing needs restart >= 2");
        double b = run();  // warm-up
        hipEvent_t e0, e1;
        NUMA_HIP_CHECK(hipEventCreate(&e0));
        NUMA_HIP_CHECK(hipEventCreate(&e1));
        NUMA_HIP_CHECK(hipEventRecord(e0, c_.stream));
        for (int64_t r = 0; r < reps; ++r) b = run();
        NUMA_HIP_CHECK(hipEventRecord(e1, c_.stream));
        NUMA_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        NUMA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (bytes) *bytes = b;
        return (double)ms / (double)reps;
    }

private:
    // misc scalar slots
    enum { S_FNORM = 0, S_XNORM, S_VNORM, S_BNORM, S_BETA, S_TMP, S_LAMBDA, S_DNORM, S_XHNORM,
           kMisc = 10 };
    // lambda as a device scalar: the kernels read it, so captured graphs stay
    // valid across set_lambda
    const double* lam_dev() { return MISC(S_LAMBDA); }
    double* H(int64_t i, int64_t j) { return sd_ + j * (m_ + 1) + i; }
    double* S(int64_t i) { return sd_ + (m_ + 1) * m_ + i; }
    double* RHS(int64_t i) { return sd_ + (m_ + 1) * m_ + (m_ + 1) + i; }
    double* MISC(int i) { return sd_ + (m_ + 1) * m_ + 2 * (m_ + 1) + i; }

    double read1(const double* dev) {
        NUMA_HIP_CHECK(hipMemcpyAsync(hs_, dev, sizeof(double), hipMemcpyDeviceToHost, c_.stream));
        c_.sync();
        return hs_[0];
    }
    void readn(const double* dev, int64_t cnt, double* host) {
        NUMA_HIP_CHECK(hipMemcpyAsync(hs_, dev, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost,
                                      c_.stream));
        c_.sync();
        std::copy(hs_, hs_ + cnt, host);
    }

    double* allocd() {
        void* p = nullptr;
        NUMA_HIP_CHECK(hipMalloc(&p, (size_t)N_ * sizeof(double)));
        allocs_.push_back(p);
        return static_cast<double*>(p);
    }
    T* alloct() {
        void* p = nullptr;
        NUMA_HIP_CHECK(hipMalloc(&p, (size_t)N_ * sizeof(T)));
        allocs_.push_back(p);
        return static_cast<T*>(p);
    }

    // ||v|| -> slot (EPI_SQRT) or ||v||, 1/||v|| -> slot, slot2
    template <class TV>
    void norm(const TV* v, double* slot, double* slot2 = nullptr) {
        launch_dot<TV, TV>(c_, N_, v, v,
                           Epilogue{slot2 ? EPI_NORM_SCALE : EPI_SQRT, 1.0, nullptr, slot, slot2});
    }

    // F at the trial point w = (x - a z) - b u (z,u may be null); writes F and,
    // if wout, w; ||F|| -> S_FNORM.
    void eval_F(const double* x, const double* z, double a, const double* u, double b, double* F,
                double* wout) {
        const Epilogue ep{EPI_SQRT, 1.0, nullptr, MISC(S_FNORM), nullptr};
        if (!z && !u)
            launch_stencil<true>(c_, n_, LdVec<double>{x}, EpF<false>{h2inv_, lambda_, F, nullptr, lam_dev()},
                                 ep);
        else if (!u)
            launch_stencil<true>(c_, n_, LdStep<double, true, false>{x, z, a, nullptr, 0.0},
                                 EpF<true>{h2inv_, lambda_, F, wout, lam_dev()}, ep);
        else
            launch_stencil<true>(c_, n_, LdStep<double, true, true>{x, z, a, u, b},
                                 EpF<true>{h2inv_, lambda_, F, wout, lam_dev()}, ep);
        if (inject_fires(NUMA_INJECT_F)) launch_poison<double>(c_, F, MISC(S_FNORM), nullptr);
    }

    // fault injection (inject_nan): true when this call of `what` is the armed one
    bool inject_fires(int64_t what) {
        if (inj_what_ != what) return false;
        if (++inj_count_ != inj_at_) return false;
        inj_what_ = -1;
        return true;
    }

    // Everything that depends only on the Newton point x: ||x|| for the FD
    // step, the Jacobi diagonal (BratuJacobi::make, bratu3d.hpp:251-255).
    void point_setup(const double* x) {
        if (E_) {  // ||x|| and lambda e^x in one pass
            const unsigned nb = stream_blocks(N_);
            if (aligned2(x, sizeof(double)))
                hipLaunchKernelGGL(k_xnorm_exp<true>, dim3(nb), dim3(kBlock), 0, c_.stream, N_, x,
                                   lambda_, lam_dev(), E_, c_.red);
            else
                hipLaunchKernelGGL(k_xnorm_exp<false>, dim3(nb), dim3(kBlock), 0, c_.stream, N_,
                                   x, lambda_, lam_dev(), E_, c_.red);
            NUMA_HIP_LAUNCH_CHECK();
            c_.finish(nb, Epilogue{EPI_SQRT, 1.0, nullptr, MISC(S_XNORM), nullptr});
        } else if (cfg_.jv_mode == NUMA_DERIV_FD || cfg_.d2_mode == NUMA_DERIV_FD) {
            norm(x, MISC(S_XNORM));
        }
        if (cfg_.precond == NUMA_PRECOND_JACOBI) {
            hipLaunchKernelGGL(k_jacobi_diag, dim3(stream_blocks(N_)), dim3(kBlock), 0, c_.stream,
                               N_, h2inv_, lambda_, lam_dev(), x, diag_);
            NUMA_HIP_LAUNCH_CHECK();
        }
    }

    // out = scale * J(x) v  [or b - J(x) v]; reduction epilogue ep if reducing.
    // Exactly one of dotv (reduce out*dotv) / b (out = b - J v, reduce out^2)
    // is set when the launch reduces.
    template <class TV>
    // vnorm_given: ||v|| is already in the S_VNORM slot (from the
    // preconditioner that produced v).
    void jv(const double* x, const double* F0, const TV* v, T* out, const double* scale,
            const T* dotv, const double* b, const Epilogue& ep, bool vnorm_given = false) {
        if (cfg_.jv_mode == NUMA_DERIV_FD && E_) {
            // mixed precision: the rearranged quotient over v and lambda e^x
            if (!vnorm_given) norm(v, MISC(S_VNORM));
            const FDStep st{std::sqrt(std::numeric_limits<double>::epsilon()), MISC(S_XNORM),
                            MISC(S_VNORM)};
            const LdVec<TV> L{v};
            if (dotv)
                launch_stencil<true>(c_, n_, L,
                                     EpJvFDe<T, true, false>{h2inv_, st, scale, E_, out, dotv,
                                                             nullptr, 0.0, 1.0},
                                     ep);
            else if (b)
                launch_stencil<true>(c_, n_, L,
                                     EpJvFDe<T, false, true>{h2inv_, st, scale, E_, out, nullptr,
                                                             b, 0.0, 1.0},
                                     ep);
            else
                launch_stencil<false>(c_, n_, L,
                                      EpJvFDe<T, false, false>{h2inv_, st, scale, E_, out, nullptr,
                                                               nullptr, 0.0, 1.0},
                                      ep);
        } else if (cfg_.jv_mode == NUMA_DERIV_FD) {
            if (!vnorm_given) norm(v, MISC(S_VNORM));
            const FDStep st{std::sqrt(std::numeric_limits<double>::epsilon()), MISC(S_XNORM),
                            MISC(S_VNORM)};
            const LdXpHV<TV> L{x, v, st, 0.0};
            if (dotv)
                launch_stencil<true>(c_, n_, L,
                                     EpJvFD<T, true, false>{h2inv_, lambda_, st, scale, F0, out,
                                                            dotv, nullptr, 0.0, 1.0, lam_dev()},
                                     ep);
            else if (b)
                launch_stencil<true>(c_, n_, L,
                                     EpJvFD<T, false, true>{h2inv_, lambda_, st, scale, F0, out,
                                                            nullptr, b, 0.0, 1.0, lam_dev()},
                                     ep);
            else
                launch_stencil<false>(c_, n_, L,
                                      EpJvFD<T, false, false>{h2inv_, lambda_, st, scale, F0, out,
                                                              nullptr, nullptr, 0.0, 1.0, lam_dev()},
                                      ep);
        } else {
            const LdVec<TV> L{v};
            if (dotv)
                launch_stencil<true>(
                    c_, n_, L,
                    EpJvAn<T, true, false>{h2inv_, lambda_, scale, x, out, dotv, nullptr, 1.0, lam_dev()}, ep);
            else if (b)
                launch_stencil<true>(
                    c_, n_, L,
                    EpJvAn<T, false, true>{h2inv_, lambda_, scale, x, out, nullptr, b, 1.0, lam_dev()}, ep);
            else
                launch_stencil<false>(
                    c_, n_, L,
                    EpJvAn<T, false, false>{h2inv_, lambda_, scale, x, out, nullptr, nullptr, 1.0, lam_dev()},
                    ep);
        }
        if (inject_fires(NUMA_INJECT_JV)) launch_poison<T>(c_, out, ep.out, ep.out2);
    }

    // Preconditioner M^-1 in -> out; returns the result pointer (in itself
    // for no preconditioner, IdentityPrecond of gmres.hpp:37-39).
    // precond(), and with vnorm (FD J v next) ||out|| into the S_VNORM slot
    // when the V-cycle can give it with its last launch; returns whether it did.
    const T* precond_for_jv(const T* in, T* out, bool& vnorm) {
        vnorm = false;
        const bool gmg = cfg_.precond == NUMA_PRECOND_GMG ||
                         (cfg_.precond == NUMA_PRECOND_LEARNED && !(net_ && use_net_));
        if (cfg_.jv_mode != NUMA_DERIV_FD || !gmg || !mg_ || capture_ || in == out)
            return precond(in, out);
        mg_->apply_norm(in, out, MISC(S_VNORM));
        if (cfg_.precond == NUMA_PRECOND_LEARNED) ++lstats_.gmg_applications;
        vnorm = true;
        return out;
    }

    const T* precond(const T* in, T* out) {
        const T* res = in;
        switch (cfg_.precond) {
            case NUMA_PRECOND_LEARNED:
                if (net_ && use_net_) {
                    net_->apply(n_, in, out);
                    ++lstats_.net_applications;
                } else {
                    mg_->apply(in, out);
                    ++lstats_.gmg_applications;
                }
                res = out;
                break;
            case NUMA_PRECOND_GMG:
                mg_->apply(in, out);
                res = out;
                break;
            case NUMA_PRECOND_JACOBI:
                launch_diag_solve<T>(c_, N_, in, diag_, out);
                res = out;
                break;
            default:
                break;
        }
        if (capture_) capture_(in, res);
        return res;
    }

    // out = J(x) v in FP64 whatever the Krylov precision; F0 = F(x) and
    // xnorm = ||x|| (device slot) are what the finite-difference form uses
    // (operator.hpp:101-114); the analytic form is bratu3d.hpp:33-43.
    void jv64(const double* x, const double* F0, const double* xnorm, const double* v,
              double* out) {
        if (cfg_.jv_mode == NUMA_DERIV_FD) {
            norm(v, MISC(S_VNORM));
            const FDStep st{std::sqrt(std::numeric_limits<double>::epsilon()), xnorm,
                            MISC(S_VNORM)};
            launch_stencil<false>(c_, n_, LdXpHV<double>{x, v, st, 0.0},
                                  EpJvFD<double, false, false>{h2inv_, lambda_, st, nullptr, F0,
                                                               out, nullptr, nullptr, 0.0, 1.0,
                                                               lam_dev()},
                                  Epilogue{});
        } else {
            launch_stencil<false>(c_, n_, LdVec<double>{v},
                                  EpJvAn<double, false, false>{h2inv_, lambda_, nullptr, x, out,
                                                               nullptr, nullptr, 1.0, lam_dev()},
                                  Epilogue{});
        }
    }

    // tensor_free_curvature_2018 (methods/tf_chebyshev_2018.hpp:23-51):
    // out = s [J(x + h d) d - J(x) d] / h with h = clamp(sqrt(eps) ||x|| / ||d||,
    // 2 eps, 0.5) (2 eps if d = 0); F0 = F(x) for the finite-difference J v
    // (computed here if null).  xh, Fh: scratch for the shifted point and its
    // F.  J v products go to tfa_ / tfb_; out may be tfa_.  Returns h.
    double tf_curvature(const double* x, const double* F0, const double* d, double* out,
                        double s, double* xh, double* Fh) {
        const double eps = std::numeric_limits<double>::epsilon();
        norm(d, MISC(S_DNORM));
        norm(x, MISC(S_XNORM));
        double nrm[2];
        readn(MISC(S_XNORM), 1, nrm);
        readn(MISC(S_DNORM), 1, nrm + 1);
        const double nd = nrm[1];
        const double h = !(nd > 0.0) ? 2.0 * eps
                                     : std::clamp(std::sqrt(eps) * nrm[0] / nd, 2.0 * eps, 0.5);
        if (!tfa_) {
            tfa_ = allocd();
            tfb_ = allocd();
        }
        launch_axpy_to(c_, N_, x, h, d, xh);  // xh = x + h d
        const bool fd = cfg_.jv_mode == NUMA_DERIV_FD;
        const double* F0x = F0;
        if (fd && !F0) {  // F(x) in Fh, consumed by the first J v before Fh = F(xh)
            eval_F(x, nullptr, 0.0, nullptr, 0.0, Fh, nullptr);
            F0x = Fh;
        }
        jv64(x, F0x, MISC(S_XNORM), d, tfa_);
        if (fd) {
            eval_F(xh, nullptr, 0.0, nullptr, 0.0, Fh, nullptr);
            norm(xh, MISC(S_XHNORM));
        }
        jv64(xh, Fh, MISC(S_XHNORM), d, tfb_);
        launch_diffq(c_, N_, tfb_, tfa_, h, s, out);
        return h;
    }

    void d2(const double* x, const double* F0, const double* v, double* out) {
        if (cfg_.d2_mode == NUMA_DERIV_FD) {
            norm(v, MISC(S_VNORM));
            const FDStep st{std::pow(std::numeric_limits<double>::epsilon(), 0.25), MISC(S_XNORM),
                            MISC(S_VNORM)};
            launch_stencil<false>(c_, n_, LdPM<double>{x, v, st, 0.0},
                                  EpD2FD{h2inv_, lambda_, st, F0, out, 0.0, lam_dev()}, Epilogue{});
        } else {
            hipLaunchKernelGGL(k_d2_analytic, dim3(stream_blocks(N_)), dim3(kBlock), 0, c_.stream,
                               N_, lambda_, lam_dev(), x, v, out);
            NUMA_HIP_LAUNCH_CHECK();
        }
    }

    numa_gmres_result gmres(const double* x0, const double* F0, const double* b, double* xl,
                            double rtol, int64_t restart, int64_t max_iters, bool zero_guess,
                            const double* r_init, bool save);
    numa_gmres_result gmres_recycled(const double* x0, const double* F0, const double* b,
                                     double* xl, double rtol, int64_t restart, int64_t max_iters,
                                     bool save = false);

    Ctx& c_;
    int64_t n_, N_;
    double lambda_, h2inv_;
    numa_solver_config cfg_;
    int64_t m_;
    std::unique_ptr<MGBase> mg_;
    // captured Arnoldi steps: (step, x0, F0) -> executable graph
    bool graphs_ = false;
    double* xa_ = nullptr;  // iterate buffer of graph mode (fixed pointers)
    std::map<std::tuple<size_t, const void*, const void*>, hipGraphExec_t> gcache_;
    void drop_graphs() {
        for (auto& kv : gcache_) (void)hipGraphExecDestroy(kv.second);
        gcache_.clear();
    }
    bool in_solve_ = false;  // graphs serve solve() (fixed iterate buffers) only
    bool graph_ok() const {
        return graphs_ && !timing_ && cfg_.precond != NUMA_PRECOND_LEARNED && !capture_ &&
               inj_what_ < 0 && !(mg_ && mg_->sweep_events_on());
    }
    int64_t inj_what_ = -1, inj_at_ = 0, inj_count_ = 0;  // inject_nan state
    // ---- timers (operators/timing.hpp:16-108): device time of each primitive
    // call, from a pair of HIP events on the solver stream, summed after the
    // solve's final synchronisation.
    enum { T_F = 0, T_JVP, T_D2, T_PREC, T_SETUP };
    bool timing_ = false;
    numa_timers tacc_{};
    struct TPair {
        int cat;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> evpool_;
    std::vector<TPair> tpend_;
    hipEvent_t ev_get() {
        if (evpool_.empty()) {
            hipEvent_t e = nullptr;
            NUMA_HIP_CHECK(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = evpool_.back();
        evpool_.pop_back();
        return e;
    }
    template <class F>
    void timed(int cat, F&& f) {
        if (!timing_) {
            f();
            return;
        }
        hipEvent_t a = ev_get(), b = ev_get();
        NUMA_HIP_CHECK(hipEventRecord(a, c_.stream));
        f();
        NUMA_HIP_CHECK(hipEventRecord(b, c_.stream));
        tpend_.push_back({cat, a, b});
    }
    void collect_timers() {  // the stream has been synchronised
        for (const TPair& t : tpend_) {
            float ms = 0.0f;
            NUMA_HIP_CHECK(hipEventElapsedTime(&ms, t.a, t.b));
            const double sec = 1e-3 * (double)ms;
            switch (t.cat) {
                case T_F: tacc_.t_F += sec; break;
                case T_JVP: tacc_.t_JVP += sec; break;
                case T_D2: tacc_.t_D2 += sec; break;
                case T_PREC: tacc_.t_prec += sec; break;
                default: tacc_.t_setup_dynamic += sec; break;
            }
            evpool_.push_back(t.a);
            evpool_.push_back(t.b);
        }
        tpend_.clear();
    }
    // Runs `body` (kernel launches only) for Arnoldi step j, from the graph
    // captured for (j, x0, F0) or by capturing it now.
    template <class F>
    void run_step(size_t j, const double* x0, const double* F0, F&& body) {
        if (!graph_ok() || !in_solve_) {
            body();
            return;
        }
        const auto key = std::make_tuple(j, (const void*)x0, (const void*)F0);
        auto it = gcache_.find(key);
        if (it == gcache_.end()) {
            if (mg_) mg_->prepare();
            NUMA_HIP_CHECK(hipStreamBeginCapture(c_.stream, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(c_.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hipGraph_t g = nullptr;
            NUMA_HIP_CHECK(hipStreamEndCapture(c_.stream, &g));
            hipGraphExec_t ge = nullptr;
            const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            NUMA_HIP_CHECK(e);
            it = gcache_.emplace(key, ge).first;
        }
        NUMA_HIP_CHECK(hipGraphLaunch(it->second, c_.stream));
    }
    std::shared_ptr<NetApply> net_;
    double ratio_ = 0.5;
    bool use_net_ = false;
    LearnedStats lstats_;
    CaptureFn capture_;
    double *x2_ = nullptr, *Fx_ = nullptr, *Fn_ = nullptr, *z_ = nullptr, *u_ = nullptr,
           *w_ = nullptr, *r_ = nullptr, *diag_ = nullptr;
    double *tfa_ = nullptr, *tfb_ = nullptr;  // J v products of the TF-2018 curvature
    float* E_ = nullptr;  // lambda e^x at the Newton point (mixed FD J v)
    std::vector<T*> V_, Z_;
    T* zb_ = nullptr;
    T* comb_ = nullptr;
    int64_t nsc_ = 0;
    double* sd_ = nullptr;
    double* hs_ = nullptr;
    std::vector<void*> allocs_;
    // saved last Arnoldi cycle for recycling (gmres.hpp:210-220)
    int64_t saved_k_ = 0;
    std::vector<std::vector<double>> saved_H_;
    std::vector<double> saved_s_;
};

// linalg/gmres.hpp:61-223 (and FGMRES when cfg.flexible).
template <class T>
numa_gmres_result BratuSolver<T>::gmres(const double* x0, const double* F0, const double* b,
                                        double* xl, double rtol, int64_t restart,
                                        int64_t max_iters, bool zero_guess, const double* r_init,
                                        bool save) {
    numa_gmres_result res{};
    if (save) saved_k_ = 0;
    // From a zero guess the first residual is b itself: ||b||, s_0 = 1/||b||
    // and V_0 = b come from one pass over b, and x_l is written (not
    // accumulated) by the first cycle's update instead of being zeroed first.
    const bool b_is_r0 = zero_guess && r_init == nullptr && restart > 0 && max_iters > 0;
    if (b_is_r0)
        launch_convert_norm<double, T>(c_, N_, b, V_[0],
                                       Epilogue{EPI_NORM_SCALE, 1.0, nullptr, MISC(S_BNORM), S(0)});
    else
        norm(b, MISC(S_BNORM));
    const double bnorm = read1(MISC(S_BNORM));
    if (bnorm == 0.0) {
        launch_fill<double>(c_, N_, 0.0, xl);
        res.converged = 1;
        return res;
    }
    if (restart == 0) {
        res.rel_res = 1.0;
        res.breakdown = 1;
        return res;
    }
    require(restart <= m_, "gmres restart exceeds the solver's basis capacity");
    if (max_iters == 0) {
        res.rel_res = 1.0;
        return res;
    }
    bool xl_unset = zero_guess;  // x_l = 0 not yet stored (see above)
    auto settle_xl = [&]() {
        if (xl_unset) launch_fill<double>(c_, N_, 0.0, xl);
        xl_unset = false;
    };
    use_net_ = net_ != nullptr;

    const bool flex = cfg_.flexible != 0;
    const size_t m = (size_t)restart;
    std::vector<std::vector<double>> Hh(m + 1, std::vector<double>(m, 0.0)),
        Hraw(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m), sn(m), g(m + 1), y(m), s(m + 1), col(m + 2);
    bool use_given = r_init != nullptr;
    bool first = zero_guess;
    size_t k_last = 0;
    const T* zlast[1] = {nullptr};
    (void)zlast;

    while (res.iters < max_iters) {
        double beta;
        if (use_given) {
            launch_convert<double, T>(c_, N_, r_init, V_[0]);
            norm(r_init, MISC(S_BETA), S(0));
            use_given = false;
            first = false;
            beta = read1(MISC(S_BETA));
        } else if (first) {
            if (b_is_r0) {
                beta = bnorm;  // V_0 and s_0 are in place
            } else {
                launch_convert<double, T>(c_, N_, b, V_[0]);
                norm(b, MISC(S_BETA), S(0));
                beta = read1(MISC(S_BETA));
            }
            first = false;
        } else {
            settle_xl();
            timed(T_JVP, [&]() {
                jv<double>(x0, F0, xl, V_[0], nullptr, nullptr, b,
                           Epilogue{EPI_NORM_SCALE, 1.0, nullptr, MISC(S_BETA), S(0)});
            });
            ++res.matvecs;
            beta = read1(MISC(S_BETA));
        }
        res.rel_res = beta / bnorm;
        if (res.rel_res <= rtol) {
            res.converged = 1;
            k_last = 0;
            break;
        }
        for (auto& row : Hraw) std::fill(row.begin(), row.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        s[0] = beta > 0.0 ? 1.0 / beta : 0.0;

        size_t k = 0;
        k_last = 0;
        while (k < m && res.iters < max_iters) {
            const size_t j = k;
            T* zdst = flex && !Z_.empty() ? Z_[j] : zb_;
            const bool by_net = cfg_.precond == NUMA_PRECOND_LEARNED && net_ && use_net_;
            run_step(j, x0, F0, [&]() {
                const T* zj = nullptr;
                bool vnorm = false;
                timed(T_PREC, [&]() { zj = precond_for_jv(V_[j], zdst, vnorm); });
                // w = s_j J z_j into V_{j+1}; first MGS dot -> H[0][j] = s_0 <w, V_0raw>
                timed(T_JVP, [&]() {
                    jv<T>(x0, F0, zj, V_[j + 1], S(j), V_[0], nullptr,
                          Epilogue{EPI_STORE, 1.0, S(0), H(0, j), nullptr}, vnorm);
                });
                for (size_t i = 0; i <= j; ++i) {
                    const bool last = i == j;
                    const Epilogue ep =
                        last ? Epilogue{EPI_NORM_SCALE, 1.0, nullptr, H(j + 1, j), S(j + 1)}
                             : Epilogue{EPI_STORE, 1.0, S(i + 1), H(i + 1, j), nullptr};
                    launch_mgs<T>(c_, N_, V_[j + 1], V_[i], H(i, j), S(i),
                                  last ? (const T*)nullptr : (const T*)V_[i + 1], ep);
                }
            });
            ++res.precs;
            ++res.matvecs;
            readn(H(0, j), (int64_t)j + 2, col.data());
            for (size_t i = 0; i <= j + 1; ++i) Hh[i][j] = col[i];
            s[j + 1] = Hh[j + 1][j] > 0.0 ? 1.0 / Hh[j + 1][j] : 0.0;
            for (size_t i = 0; i <= j + 1; ++i) Hraw[i][j] = Hh[i][j];
            for (size_t i = 0; i < j; ++i) {
                const double t = cs[i] * Hh[i][j] + sn[i] * Hh[i + 1][j];
                Hh[i + 1][j] = -sn[i] * Hh[i][j] + cs[i] * Hh[i + 1][j];
                Hh[i][j] = t;
            }
            const double denom = std::hypot(Hh[j][j], Hh[j + 1][j]);
            if (!(denom > 0.0) || !std::isfinite(denom)) {
                res.breakdown = 1;
                settle_xl();
                return res;
            }
            cs[j] = Hh[j][j] / denom;
            sn[j] = Hh[j + 1][j] / denom;
            Hh[j][j] = denom;
            Hh[j + 1][j] = 0.0;
            const double gprev = std::abs(g[j]);
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            if (by_net && std::abs(g[j + 1]) > ratio_ * gprev) {
                use_net_ = false;  // the V-cycle serves the rest of this solve
                ++lstats_.fallbacks;
            }
            ++k;
            ++res.iters;
            k_last = k;
            res.rel_res = std::abs(g[j + 1]) / bnorm;
            if (res.rel_res <= rtol) break;
        }
        for (size_t ii = k; ii-- > 0;) {
            const double pivot = Hh[ii][ii];
            if (!(std::abs(pivot) > 0.0) || !std::isfinite(pivot)) {
                res.breakdown = 1;
                settle_xl();
                return res;
            }
            double acc = g[ii];
            for (size_t jj = ii + 1; jj < k; ++jj) acc -= Hh[ii][jj] * y[jj];
            y[ii] = acc / pivot;
        }
        std::vector<const T*> vecs(k);
        std::vector<double> coef(k);
        for (size_t i = 0; i < k; ++i) coef[i] = y[i] * s[i];
        if (flex) {
            for (size_t i = 0; i < k; ++i) vecs[i] = Z_.empty() ? V_[i] : Z_[i];
            launch_combine<T, double>(c_, N_, vecs, coef, xl, !xl_unset);  // 0 + sum: same bits
        } else {
            for (size_t i = 0; i < k; ++i) vecs[i] = V_[i];
            launch_combine<T, T>(c_, N_, vecs, coef, comb_, false);
            const T* zq = nullptr;
            timed(T_PREC, [&]() { zq = precond(comb_, zb_); });
            ++res.precs;
            if (xl_unset)
                launch_convert<T, double>(c_, N_, zq, xl);  // 0 + 1 * zq
            else
                launch_axpy<T, double>(c_, N_, 1.0, zq, xl);
        }
        xl_unset = false;
        if (res.rel_res <= rtol) {
            res.converged = 1;
            break;
        }
    }
    settle_xl();
    if (save && k_last > 0) {
        saved_k_ = (int64_t)k_last;
        saved_H_.assign(k_last + 1, std::vector<double>(k_last, 0.0));
        for (size_t i = 0; i <= k_last; ++i)
            for (size_t j = 0; j < k_last; ++j) saved_H_[i][j] = Hraw[i][j];
        saved_s_.assign(s.begin(), s.begin() + k_last + 1);
    }
    return res;
}

// Second solve on the saved Arnoldi relation (gmres.hpp:229-323).  The saved
// basis is still resident in V; it is consumed before GMRES overwrites it.
template <class T>
numa_gmres_result BratuSolver<T>::gmres_recycled(const double* x0, const double* F0,
                                                 const double* b, double* xl, double rtol,
                                                 int64_t restart, int64_t max_iters, bool save) {
    launch_fill<double>(c_, N_, 0.0, xl);
    const int64_t k = saved_k_;
    if (k == 0) return gmres(x0, F0, b, xl, rtol, restart, max_iters, true, nullptr, save);
    std::vector<double> rhs((size_t)k + 1);
    for (int64_t i = 0; i <= k; ++i)
        launch_dot<T, double>(c_, N_, V_[(size_t)i], b,
                              Epilogue{EPI_STORE, 1.0, S(i), RHS(i), nullptr});
    readn(RHS(0), k + 1, rhs.data());
    auto R = saved_H_;
    std::vector<double> cs((size_t)k, 0.0), sn((size_t)k, 0.0);
    for (int64_t j = 0; j < k; ++j) {
        for (int64_t i = 0; i < j; ++i) {
            const double t = cs[i] * R[i][j] + sn[i] * R[i + 1][j];
            R[i + 1][j] = -sn[i] * R[i][j] + cs[i] * R[i + 1][j];
            R[i][j] = t;
        }
        const double denom = std::hypot(R[j][j], R[j + 1][j]);
        if (!(denom > 0.0) || !std::isfinite(denom))
            return gmres(x0, F0, b, xl, rtol, restart, max_iters, true, nullptr, save);
        cs[j] = R[j][j] / denom;
        sn[j] = R[j + 1][j] / denom;
        R[j][j] = denom;
        R[j + 1][j] = 0.0;
        const double t = cs[j] * rhs[j] + sn[j] * rhs[j + 1];
        rhs[j + 1] = -sn[j] * rhs[j] + cs[j] * rhs[j + 1];
        rhs[j] = t;
    }
    std::vector<double> y((size_t)k, 0.0);
    for (int64_t ii = k; ii-- > 0;) {
        const double pivot = R[ii][ii];
        if (!(std::abs(pivot) > 0.0) || !std::isfinite(pivot)) {
            launch_fill<double>(c_, N_, 0.0, xl);
            return gmres(x0, F0, b, xl, rtol, restart, max_iters, true, nullptr, save);
        }
        double acc = rhs[ii];
        for (int64_t jj = ii + 1; jj < k; ++jj) acc -= R[ii][jj] * y[jj];
        y[ii] = acc / pivot;
    }
    // residual = b - V_{k+1} Hbar y  (no A v products)
    std::vector<const T*> vecs((size_t)k + 1);
    std::vector<double> coef((size_t)k + 1);
    for (int64_t i = 0; i <= k; ++i) {
        double cf = 0.0;
        for (int64_t j = 0; j < k; ++j) cf += saved_H_[i][j] * y[j];
        vecs[i] = V_[(size_t)i];
        coef[i] = -cf * saved_s_[i];
    }
    launch_convert<double, double>(c_, N_, b, r_);
    launch_combine<T, double>(c_, N_, vecs, coef, r_, true);
    // x0 = M^-1 (V_k y)
    std::vector<const T*> vk((size_t)k);
    std::vector<double> ck((size_t)k);
    for (int64_t j = 0; j < k; ++j) {
        vk[j] = V_[(size_t)j];
        ck[j] = y[j] * saved_s_[j];
    }
    launch_combine<T, T>(c_, N_, vk, ck, comb_, false);
    const T* px = nullptr;
    timed(T_PREC, [&]() { px = precond(comb_, zb_); });
    launch_convert<T, double>(c_, N_, px, xl);
    auto result = gmres(x0, F0, b, xl, rtol, restart, max_iters, false, r_, save);
    ++result.precs;
    return result;
}

// methods/krylov_newton.hpp:69-189 (Newton when method = NEWTON, inexact
// Chebyshev with optional recycling when method = CHEBYSHEV).
template <class T>
void BratuSolver<T>::solve(const numa_krylov_params& p, double* xdev, numa_krylov_result& out) {
    const auto t0 = std::chrono::steady_clock::now();
    out = numa_krylov_result{};
    const bool cheb = cfg_.method == NUMA_METHOD_CHEBYSHEV;
    const bool tf = cfg_.method == NUMA_METHOD_TF2018;
    const bool rec = cheb && p.recycle;
    require(p.gmres_restart >= 0 && p.gmres_restart <= m_,
            "gmres_restart exceeds the solver's basis capacity");
    // graph mode iterates in the solver's own buffers (fixed graph arguments)
    const bool own = graph_ok();
    if (own) launch_convert<double, double>(c_, N_, xdev, xa_);
    struct InSolve {
        bool& f;
        ~InSolve() { f = false; }
    } in_solve_guard{in_solve_};
    in_solve_ = own;
    double* xc = own ? xa_ : xdev;  // current iterate
    double* xt = x2_;               // trial point
    double* Fc = Fx_;
    double* Ft = Fn_;
    timed(T_F, [&]() { eval_F(xc, nullptr, 0.0, nullptr, 0.0, Fc, nullptr); });
    ++out.n_F;
    double rnorm = read1(MISC(S_FNORM));
    out.res0 = rnorm;
    double rprev = rnorm, eta_prev = p.eta_init;
    // tensor-free Chebyshev 2018 (methods/tf_chebyshev_2018.hpp:53-132):
    // J z = F (eta1), d1 = -z, J d2 = -C d1 / 2 (eta2), x+ = (x + d1) + d2
    const double eps = std::numeric_limits<double>::epsilon();
    for (int64_t k = 0; tf && k < p.max_nl; ++k) {
        if (rnorm <= p.rtol_nl * out.res0) {
            out.converged = 1;
            break;
        }
        const double eta1 = std::clamp(std::min(1e-8, rnorm * rnorm), 2.0 * eps, 1e-8);
        const double eta2 = std::clamp(std::min(1e-8, rnorm), 2.0 * eps, 1e-8);
        timed(T_SETUP, [&]() { point_setup(xc); });
        const auto g1 = gmres(xc, Fc, Fc, z_, eta1, p.gmres_restart, p.gmres_max_iters, true,
                              nullptr, false);
        out.n_JVP += g1.matvecs;
        out.n_lin_iters += g1.iters;
        out.n_prec += g1.precs;
        if (!g1.converged) {
            ++out.n_lin_failures;
            break;
        }
        if (g1.breakdown) break;
        launch_scale_to(c_, N_, -1.0, z_, w_);  // d1
        timed(T_JVP, [&]() { tf_curvature(xc, Fc, w_, tfa_, -0.5, xt, Ft); });
        out.n_JVP += 2;
        const auto g2 = gmres(xc, Fc, tfa_, u_, eta2, p.gmres_restart, p.gmres_max_iters, true,
                              nullptr, false);
        out.n_JVP += g2.matvecs;
        out.n_lin_iters += g2.iters;
        out.n_prec += g2.precs;
        if (!g2.converged) {
            ++out.n_lin_failures;
            break;
        }
        if (g2.breakdown) break;
        // (x - 1 z) - (-1) d2 = (x + d1) + d2, and F there
        timed(T_F, [&]() { eval_F(xc, z_, 1.0, u_, -1.0, Ft, xt); });
        ++out.n_F;
        rnorm = read1(MISC(S_FNORM));
        std::swap(xc, xt);
        std::swap(Fc, Ft);
        ++out.iters;
        if (!std::isfinite(rnorm)) break;
    }
    for (int64_t k = 0; !tf && k < p.max_nl; ++k) {
        if (nl_converged(p, rnorm, out.res0, out.iters, rprev)) {
            out.converged = 1;
            break;
        }
        const double eta1 = forcing(p, k, rnorm, rprev, eta_prev, out.res0);
        const double eta2 = std::min(p.eta2_max, p.eta2_factor * eta1);
        timed(T_SETUP, [&]() { point_setup(xc); });
        // z = 0 before the solve (krylov_newton.hpp:88, 136); GMRES zeroes it
        // itself unless it returns before iterating
        if (p.gmres_restart == 0 || p.gmres_max_iters == 0) launch_fill<double>(c_, N_, 0.0, z_);
        const auto g1 = gmres(xc, Fc, Fc, z_, eta1, p.gmres_restart, p.gmres_max_iters, true,
                              nullptr, rec);
        out.n_JVP += g1.matvecs;
        out.n_lin_iters += g1.iters;
        out.n_prec += g1.precs;
        if (!g1.converged) ++out.n_lin_failures;
        if (g1.breakdown) break;
        bool ok = false;
        double fn = 0.0;
        if (cheb) {
            timed(T_D2, [&]() { d2(xc, Fc, z_, w_); });
            ++out.n_D2;
            if (p.gmres_restart == 0 || p.gmres_max_iters == 0) launch_fill<double>(c_, N_, 0.0, u_);
            const auto g2 = rec ? gmres_recycled(xc, Fc, w_, u_, eta2, p.gmres_restart,
                                                 p.gmres_max_iters)
                                : gmres(xc, Fc, w_, u_, eta2, p.gmres_restart, p.gmres_max_iters,
                                        true, nullptr, false);
            out.n_JVP += g2.matvecs;
            out.n_lin_iters += g2.iters;
            out.n_prec += g2.precs;
            if (!g2.converged) ++out.n_lin_failures;
            if (!g2.breakdown) {
                timed(T_F, [&]() { eval_F(xc, z_, 1.0, u_, 0.5, Ft, xt); });
                ++out.n_F;
                fn = read1(MISC(S_FNORM));
                if (fn < rnorm) ok = true;
            }
        }
        if (!ok) {
            double lam = 1.0;
            while (lam >= p.min_lambda) {
                timed(T_F, [&]() { eval_F(xc, z_, lam, nullptr, 0.0, Ft, xt); });
                ++out.n_F;
                fn = read1(MISC(S_FNORM));
                if (fn <= (1.0 - p.armijo * lam) * rnorm) {
                    ok = true;
                    break;
                }
                lam *= 0.5;
            }
        }
        if (!ok) break;
        std::swap(xc, xt);
        std::swap(Fc, Ft);
        rprev = rnorm;
        rnorm = fn;
        eta_prev = eta1;
        ++out.iters;
    }
    if (rnorm <= p.rtol_nl * out.res0) out.converged = 1;
    out.res = rnorm;
    if (own) {
        launch_convert<double, double>(c_, N_, xc, xdev);
        xa_ = xc;  // the pair {xa_, x2_} is kept, in either order
        x2_ = xt;
    } else if (xc != xdev) {
        launch_convert<double, double>(c_, N_, xc, xdev);
        x2_ = xc;  // the internal buffer now holds the previous trial storage
    }
    // keep Fx_/Fn_ ownership consistent for the next call
    Fx_ = Fc;
    Fn_ = Ft;
    c_.sync();
    collect_timers();
    out.t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

SolverBase* make_bratu_solver(Ctx& c, int64_t n, double lambda, const numa_solver_config& cfg,
                              int64_t restart) {
    if (cfg.precision == NUMA_PRECISION_MIXED)
        return new BratuSolver<float>(c, n, lambda, cfg, restart);
    if (cfg.precision == NUMA_PRECISION_FP64)
        return new BratuSolver<double>(c, n, lambda, cfg, restart);
    throw Error(NUMA_ERR_ARG, "unknown precision");
}

}  // namespace numa_hip

// ===========================================================================
// C entry points: multigrid and solver handles
// ===========================================================================
using numa_hip::Error;


extern "C" {

int numa_hip_mg_create(numa_hip_ctx* ctx, int64_t n, const numa_mg_config* cfg,
                       int64_t precision, numa_hip_mg** out) {
    NUMA_API_BEGIN
    numa_hip::require(ctx && cfg && out, "null argument");
    numa_hip::require(precision == NUMA_PRECISION_FP64 || precision == NUMA_PRECISION_MIXED,
                      "unknown precision");
    NUMA_HIP_CHECK(hipSetDevice(ctx->impl->device));
    auto h = std::make_unique<numa_hip_mg>();
    h->ctx = ctx->impl;
    h->precision = precision;
    h->impl.reset(numa_hip::make_mg(*ctx->impl, n, *cfg, precision));
    ++ctx->impl->live_handles;
    *out = h.release();
    NUMA_API_END
}

int numa_hip_mg_levels(const numa_hip_mg* mg, int64_t* levels, int64_t* sizes_host) {
    NUMA_API_BEGIN
    numa_hip::require(mg && levels, "null argument");
    *levels = mg->impl->levels();
    if (sizes_host)
        for (int64_t l = 0; l < *levels; ++l) sizes_host[l] = mg->impl->level_n(l);
    NUMA_API_END
}

int numa_hip_mg_apply(numa_hip_mg* mg, const void* r_dev, void* z_dev) {
    NUMA_API_BEGIN
    numa_hip::require(mg && r_dev && z_dev, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(mg->ctx->device));
    mg->impl->apply(r_dev, z_dev);
    NUMA_API_END
}

int numa_hip_mg_apply_norm(numa_hip_mg* mg, const void* r_dev, void* z_dev, double* norm_host) {
    NUMA_API_BEGIN
    numa_hip::require(mg && r_dev && z_dev && norm_host, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(mg->ctx->device));
    numa_hip::Ctx& c = *mg->ctx;
    double* slot = c.dscal + (numa_hip::Ctx::kScalars - 1);  // this entry's own slot
    mg->impl->apply_norm(r_dev, z_dev, slot);
    c.read_scalars(slot, norm_host, 1);
    NUMA_API_END
}

int numa_hip_mg_destroy(numa_hip_mg* mg) {
    NUMA_API_BEGIN
    if (mg) {
        NUMA_HIP_CHECK(hipStreamSynchronize(mg->ctx->stream));
        --mg->ctx->live_handles;
        delete mg;
    }
    NUMA_API_END
}

int numa_hip_bratu_solver_create(numa_hip_ctx* ctx, int64_t n, double lambda,
                                 const numa_solver_config* cfg, int64_t gmres_restart,
                                 numa_hip_solver** out) {
    NUMA_API_BEGIN
    numa_hip::require(ctx && cfg && out, "null argument");
    numa_hip::require(std::isfinite(lambda), "lambda must be finite");
    NUMA_HIP_CHECK(hipSetDevice(ctx->impl->device));
    auto h = std::make_unique<numa_hip_solver>();
    h->ctx = ctx->impl;
    h->impl.reset(numa_hip::make_bratu_solver(*ctx->impl, n, lambda, *cfg, gmres_restart));
    ++ctx->impl->live_handles;
    *out = h.release();
    NUMA_API_END
}

int numa_hip_solver_destroy(numa_hip_solver* s) {
    NUMA_API_BEGIN
    if (s) {
        NUMA_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        --s->ctx->live_handles;
        delete s;
    }
    NUMA_API_END
}

int numa_hip_gmres(numa_hip_solver* s, const double* x0_dev, const double* b_dev, double* z_dev,
                   double rtol, int64_t restart, int64_t max_iters, int64_t zero_guess,
                   numa_gmres_result* out) {
    NUMA_API_BEGIN
    numa_hip::require(s && x0_dev && b_dev && z_dev && out, "null argument");
    numa_hip::require(restart >= 0 && max_iters >= 0, "negative restart / max_iters");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    *out = s->impl->gmres_entry(x0_dev, b_dev, z_dev, rtol, restart, max_iters, zero_guess != 0);
    NUMA_API_END
}

int numa_hip_gmres_ex(numa_hip_solver* s, const double* x0_dev, const double* b_dev,
                      double* z_dev, double rtol, int64_t restart, int64_t max_iters,
                      int64_t zero_guess, const double* r_init_dev, int64_t save,
                      numa_gmres_result* out) {
    NUMA_API_BEGIN
    numa_hip::require(s && x0_dev && b_dev && z_dev && out, "null argument");
    numa_hip::require(restart >= 0 && max_iters >= 0, "negative restart / max_iters");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    *out = s->impl->gmres_entry(x0_dev, b_dev, z_dev, rtol, restart, max_iters, zero_guess != 0,
                                r_init_dev, save != 0);
    NUMA_API_END
}

int numa_hip_gmres_space(numa_hip_solver* s, int64_t* k, double* hbar_host, double* const* V_dev) {
    NUMA_API_BEGIN
    numa_hip::require(s && k, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    *k = s->impl->saved_space(hbar_host, V_dev);
    NUMA_API_END
}

int numa_hip_gmres_recycled(numa_hip_solver* s, const double* x0_dev, const double* b_dev,
                            double* z_dev, double rtol, int64_t restart, int64_t max_iters,
                            int64_t k, const double* hbar_host, const double* const* V_dev,
                            int64_t save, numa_gmres_result* out) {
    NUMA_API_BEGIN
    numa_hip::require(s && x0_dev && b_dev && z_dev && out, "null argument");
    numa_hip::require(restart >= 0 && max_iters >= 0, "negative restart / max_iters");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    *out = s->impl->gmres_recycled_entry(x0_dev, b_dev, z_dev, rtol, restart, max_iters, k,
                                         hbar_host, V_dev, save != 0);
    NUMA_API_END
}

int numa_hip_solve(numa_hip_solver* s, const numa_krylov_params* p, double* x_dev,
                   numa_krylov_result* out) {
    NUMA_API_BEGIN
    numa_hip::require(s && p && x_dev && out, "null argument");
    numa_hip::require(p->max_nl >= 0 && p->gmres_max_iters >= 0 && p->gmres_restart >= 0,
                      "negative iteration limits");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    s->impl->solve(*p, x_dev, *out);
    NUMA_API_END
}

int numa_hip_primitive_cost(numa_hip_solver* s, int64_t prim, int64_t reps,
                            double* ms_per_call, double* bytes_per_call) {
    NUMA_API_BEGIN
    numa_hip::require(s && ms_per_call, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    *ms_per_call = s->impl->primitive_cost(prim, reps, bytes_per_call);
    NUMA_API_END
}

int numa_hip_mg_fusion(const numa_hip_mg* mg, int64_t* mode_host, int64_t* ichunk_host,
                       int64_t* rows_host) {
    NUMA_API_BEGIN
    numa_hip::require(mg && mode_host && ichunk_host && rows_host, "null argument");
    for (int64_t l = 0; l < mg->impl->levels(); ++l)
        mg->impl->fusion(l, mode_host + l, ichunk_host + l, rows_host + l);
    NUMA_API_END
}

int numa_hip_mg_set_resid_restrict(numa_hip_mg* mg, int64_t ich) {
    NUMA_API_BEGIN
    numa_hip::require(mg != nullptr, "null argument");
    mg->impl->set_resid_restrict(ich);
    NUMA_API_END
}

int numa_hip_solver_mg_fusion(const numa_hip_solver* s, int64_t* levels, int64_t* mode_host,
                              int64_t* ichunk_host, int64_t* rows_host) {
    NUMA_API_BEGIN
    numa_hip::require(s && levels, "null argument");
    const numa_hip::MGBase* mg = s->impl->multigrid();
    *levels = mg ? mg->levels() : 0;
    if (mg && mode_host && ichunk_host && rows_host)
        for (int64_t l = 0; l < mg->levels(); ++l)
            mg->fusion(l, mode_host + l, ichunk_host + l, rows_host + l);
    NUMA_API_END
}

int numa_hip_solver_sweep_events(numa_hip_solver* s, int64_t enable) {
    NUMA_API_BEGIN
    numa_hip::require(s != nullptr, "null argument");
    numa_hip::MGBase* mg = s->impl->multigrid();
    numa_hip::require(mg != nullptr, "the solver has no multigrid preconditioner");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    mg->set_sweep_events(enable != 0);
    NUMA_API_END
}

int numa_hip_solver_sweep_time(numa_hip_solver* s, double* avg_ms, int64_t* count) {
    NUMA_API_BEGIN
    numa_hip::require(s && avg_ms && count, "null argument");
    numa_hip::MGBase* mg = s->impl->multigrid();
    numa_hip::require(mg != nullptr, "the solver has no multigrid preconditioner");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    mg->sweep_time(avg_ms, count);
    NUMA_API_END
}

int numa_hip_eigen_smallest(numa_hip_solver* s, const double* x_dev, const numa_eigen_params* p,
                            numa_eigen_result* out, double* vec_dev) {
    NUMA_API_BEGIN
    numa_hip::require(s && x_dev && p && out, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    *out = s->impl->eigen_smallest(x_dev, *p, vec_dev);
    NUMA_API_END
}

int numa_hip_tf2018_curvature(numa_hip_solver* s, const double* x_dev, const double* d_dev,
                              double* out_dev, double* h_used) {
    NUMA_API_BEGIN
    numa_hip::require(s && x_dev && d_dev && out_dev, "null argument");
    numa_hip::require(out_dev != x_dev && out_dev != d_dev, "out must not alias x or d");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    const double h = s->impl->tf2018_curvature(x_dev, d_dev, out_dev);
    if (h_used) *h_used = h;
    NUMA_API_END
}

int numa_hip_solver_set_graphs(numa_hip_solver* s, int64_t enable) {
    NUMA_API_BEGIN
    numa_hip::require(s != nullptr, "null solver");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    NUMA_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    s->impl->set_graphs(enable != 0);
    NUMA_API_END
}

int numa_hip_solver_graphs_cached(const numa_hip_solver* s, int64_t* count) {
    NUMA_API_BEGIN
    numa_hip::require(s && count, "null argument");
    *count = s->impl->graphs_cached();
    NUMA_API_END
}

int numa_hip_solver_set_timing(numa_hip_solver* s, int64_t enable) {
    NUMA_API_BEGIN
    numa_hip::require(s != nullptr, "null solver");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    s->impl->set_timing(enable != 0);
    NUMA_API_END
}

int numa_hip_solver_timers(numa_hip_solver* s, numa_timers* out, int64_t reset) {
    NUMA_API_BEGIN
    numa_hip::require(s && out, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    *out = s->impl->timers(reset != 0);
    NUMA_API_END
}

int numa_hip_solver_inject_nan(numa_hip_solver* s, int64_t what, int64_t call) {
    NUMA_API_BEGIN
    numa_hip::require(s != nullptr, "null solver");
    s->impl->inject_nan(what, call);
    NUMA_API_END
}

int numa_hip_solver_learned_stats(const numa_hip_solver* s, int64_t* stats3) {
    NUMA_API_BEGIN
    numa_hip::require(s && stats3, "null argument");
    const numa_hip::LearnedStats st = s->impl->learned_stats();
    stats3[0] = st.net_applications;
    stats3[1] = st.gmg_applications;
    stats3[2] = st.fallbacks;
    NUMA_API_END
}

int numa_hip_solve_host(numa_hip_solver* s, const numa_krylov_params* p, double* x_host,
                        numa_krylov_result* out) {
    NUMA_API_BEGIN
    numa_hip::require(s && p && x_host && out, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(s->ctx->device));
    const size_t bytes = (size_t)s->impl->size() * sizeof(double);
    void* d = nullptr;
    NUMA_HIP_CHECK(hipMalloc(&d, bytes));
    struct Guard {
        void* p;
        ~Guard() { (void)hipFree(p); }
    } guard{d};
    NUMA_HIP_CHECK(hipMemcpyAsync(d, x_host, bytes, hipMemcpyHostToDevice, s->ctx->stream));
    numa_hip::require(p->max_nl >= 0, "negative iteration limits");
    s->impl->solve(*p, static_cast<double*>(d), *out);
    NUMA_HIP_CHECK(hipMemcpyAsync(x_host, d, bytes, hipMemcpyDeviceToHost, s->ctx->stream));
    NUMA_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    NUMA_API_END
}

}  // extern "C"
