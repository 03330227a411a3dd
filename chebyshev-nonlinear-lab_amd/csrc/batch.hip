// batch.hip - many Bratu3D instances (different grid size n and lambda)
// solved together: mixed-precision Newton-FGMRES(m) with the multigrid
// V-cycle, every operation one kernel launch over the instance table
// (batch.hpp), every Arnoldi step one hipGraph launch for all instances.
//
// The control flow of each instance is the one of the single-instance solver
// (BratuSolver<float>::solve / gmres, methods/krylov_newton.hpp:69-114,
// linalg/gmres.hpp:61-223): the host keeps every instance's Newton and GMRES
// state and makes its decisions from one batched read of the scalars all
// instances need (||F||, the Hessenberg column, ...).  The instances advance
// in lockstep: all instances in a GMRES cycle take Arnoldi step j together,
// an instance that has converged (or finished its cycle, or broken down) is
// switched off by its device flag and the others go on.  The V-cycle runs its
// smoothing steps one launch each (the unfused form, bit for bit the fused
// one), so each instance's iterates are bit for bit those of the
// single-instance solver with the same configuration and unfused smoothing.
#include <chrono>
#include <cmath>
#include <functional>
#include <limits>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <sstream>

#include "batch.hpp"
#include "tune.hpp"

namespace numa_hip {

class BatchSolver {
public:
    BatchSolver(Ctx& c, int64_t count, const int64_t* n, const double* lambda,
                const numa_solver_config& cfg, int64_t restart)
        : c_(c), cfg_(cfg), m_(restart), arena_(c.stream) {
        require(count >= 1 && count <= kBatchMax,
                "a batch holds 1 to " + std::to_string(kBatchMax) + " instances");
        require(restart >= 1 && restart <= kMaxComb, "batch gmres restart must be in [1, 32]");
        require(cfg.method == NUMA_METHOD_NEWTON, "the batched solver runs Newton's method");
        require(cfg.precision == NUMA_PRECISION_MIXED, "the batched solver runs mixed precision");
        require(cfg.precond == NUMA_PRECOND_GMG && cfg.flexible,
                "the batched solver runs FGMRES with the multigrid V-cycle");
        require(cfg.jv_mode == NUMA_DERIV_FD, "the batched solver uses the FD J v");
        const numa_mg_config& g = cfg.mg;
        require(g.smoother == NUMA_SMOOTH_CHEBYSHEV || g.smoother == NUMA_SMOOTH_WJACOBI,
                "the batched V-cycle smooths with Chebyshev or weighted Jacobi");
        require(g.pre >= 0 && g.post >= 0, "multigrid needs non-negative sweep counts");
        require(g.smoother != NUMA_SMOOTH_CHEBYSHEV || (g.degree >= 1 && g.cheb_ratio > 1.0),
                "Chebyshev smoother needs degree >= 1 and cheb_ratio > 1");
        require(g.coarse_iters < 0, "the batched V-cycle solves the coarsest level exactly");
        B_ = (size_t)count;
        inst_.resize(B_);
        nsc_ = (m_ + 1) * m_ + (m_ + 1) + m_ + kMisc;
        void* p = nullptr;
        NUMA_HIP_CHECK(hipMalloc(&p, B_ * (size_t)nsc_ * sizeof(double)));
        allocs_.push_back(p);
        sd_ = static_cast<double*>(p);
        NUMA_HIP_CHECK(hipMalloc(&p, B_ * 2 * sizeof(int)));
        allocs_.push_back(p);
        ctl_ = static_cast<int*>(p);
        NUMA_HIP_CHECK(hipMalloc(&p, B_ * sizeof(int)));
        allocs_.push_back(p);
        act_d_ = static_cast<int*>(p);
        NUMA_HIP_CHECK(hipHostMalloc(&p, B_ * (size_t)nsc_ * sizeof(double), hipHostMallocDefault));
        hs_ = static_cast<double*>(p);
        NUMA_HIP_CHECK(hipHostMalloc(&p, B_ * (size_t)(m_ + 2) * sizeof(double),
                                     hipHostMallocDefault));
        hup_ = static_cast<double*>(p);
        NUMA_HIP_CHECK(hipHostMalloc(&p, B_ * 2 * sizeof(int), hipHostMallocDefault));
        hctl_ = static_cast<int*>(p);
        NUMA_HIP_CHECK(hipHostMalloc(&p, B_ * sizeof(int), hipHostMallocDefault));
        hact_ = static_cast<int*>(p);
        act_.assign(B_, -1);
        size_t levels = 0;
        for (size_t i = 0; i < B_; ++i) {
            require(n[i] >= 2 && n[i] <= kBatchMaxN,
                    "batch instance grid size must be in [2, " + std::to_string(kBatchMaxN) + "]");
            require(std::isfinite(lambda[i]), "lambda must be finite");
            setup(inst_[i], n[i], lambda[i]);
            if (i == 0) levels = inst_[i].lv.size();
            require(inst_[i].lv.size() == levels,
                    "every instance of a batch needs the same number of multigrid levels");
        }
        // lambda slots (the kernels read lambda from the device)
        for (size_t i = 0; i < B_; ++i) hup_[i] = inst_[i].lambda;
        NUMA_HIP_CHECK(hipMemcpy2DAsync(MISC(inst_[0], S_LAMBDA), nsc_ * sizeof(double), hup_,
                                        sizeof(double), sizeof(double), B_,
                                        hipMemcpyHostToDevice, c_.stream));
        c_.sync();
        lvl_box_.assign(levels, 0);
        lvl_rr_.assign(levels, 0);
        gst_.resize(B_);
        for (auto& s : gst_) {
            s.Hh.assign(m_ + 1, std::vector<double>(m_, 0.0));
            s.Hraw.assign(m_ + 1, std::vector<double>(m_, 0.0));
            s.cs.assign(m_, 0.0);
            s.sn.assign(m_, 0.0);
            s.g.assign(m_ + 1, 0.0);
            s.y.assign(m_, 0.0);
            s.s.assign(m_ + 1, 0.0);
            s.col.assign(m_ + 2, 0.0);
        }
    }
    ~BatchSolver() {
        drop_graphs();
        (void)hipStreamSynchronize(c_.stream);
        for (void* p : allocs_) (void)hipFree(p);
        if (hs_) (void)hipHostFree(hs_);
        if (hup_) (void)hipHostFree(hup_);
        if (hctl_) (void)hipHostFree(hctl_);
        if (hact_) (void)hipHostFree(hact_);
    }

    void set_graphs(bool on) {
        if (!on) drop_graphs();
        graphs_ = on;
    }
    // Smoothing form of levels 0 .. count-1 (0: one launch per step, s in
    // 1..3: box shape s); the other levels keep theirs, the autotuner is off.
    void set_smoothing(int64_t count, const int64_t* form) {
        const size_t nl = inst_[0].lv.size();
        if (lvl_box_.size() != nl) lvl_box_.assign(nl, 0);
        if (lvl_rr_.size() != nl) lvl_rr_.assign(nl, 0);
        require(count >= 0 && (count == 0 || form), "bad smoothing forms");
        for (int64_t l = 0; l < count && (size_t)l + 1 < nl; ++l) {
            require(form[l] >= 0 && form[l] <= kSweepForms, "smoothing form must be in [0, 7]");
            require(form[l] == 0 || box_ok(), "box smoothing needs the Chebyshev smoother of degree <= 3");
            lvl_box_[(size_t)l] = (int)form[l];
        }
        tuned_ = true;
        drop_graphs();
    }

    void solve(const numa_krylov_params& p, double* const* x, numa_krylov_result* out);

    numa_batch_stats stats() const {
        numa_batch_stats s{};
        s.launches = launches_;
        s.graph_launches = graph_launches_;
        s.arnoldi_steps = steps_;
        s.host_syncs = syncs_;
        s.bytes = bytes_;
        s.tables = (int64_t)arena_.tables();
        s.table_bytes = (int64_t)arena_.bytes();
        s.graphs = (int64_t)gcache_.size();
        return s;
    }
    void reset_stats() {
        launches_ = graph_launches_ = steps_ = syncs_ = 0;
        bytes_ = 0.0;
    }

private:
    static constexpr int64_t kBatchMaxN = 127;        // launch_stencil's forms below 128^3
    static constexpr int kSweepForms = 7;             // smoothing forms 0 .. 7 (set_smoothing)
    static constexpr int64_t kBatchPartials = 16384;  // per-instance reduction partials
    enum { S_FNORM = 0, S_XNORM, S_VNORM, S_BNORM, S_BETA, S_LAMBDA, S_LSA, kMisc = 8 };

    struct Level {
        int64_t n = 0, N = 0;
        double h2inv = 0.0;
        float *x = nullptr, *x2 = nullptr, *b = nullptr, *d = nullptr;
        std::vector<double> c1, c2;
        int rr_ich = 0;
    };
    struct Inst {
        int64_t n = 0, N = 0;
        double lambda = 0.0, h2inv = 0.0;
        double *xa = nullptr, *x2 = nullptr, *Fx = nullptr, *Fn = nullptr, *z = nullptr;
        float* E = nullptr;
        std::vector<float*> V, Z;
        double* sd = nullptr;
        int* ctl = nullptr;
        double* partials = nullptr;
        double* dst_tab = nullptr;
        std::vector<Level> lv;
        double *xc = nullptr, *xt = nullptr, *Fc = nullptr, *Ft = nullptr;  // solve state
        float* vb = nullptr;    // V-cycle input of the step being built
        float* vout = nullptr;  // V-cycle output of the step being built
    };
    // GMRES state of one instance (gmres.hpp:61-223)
    struct GState {
        numa_gmres_result res{};
        bool in = false, first = true, xl_unset = true, cyc = false;
        double bnorm = 0.0, rtol = 0.0, beta = 0.0;
        size_t k = 0;
        std::vector<std::vector<double>> Hh, Hraw;
        std::vector<double> cs, sn, g, y, s, col;
    };

    double* H(Inst& I, int64_t i, int64_t j) { return I.sd + j * (m_ + 1) + i; }
    double* S(Inst& I, int64_t i) { return I.sd + (m_ + 1) * m_ + i; }
    double* COEF(Inst& I, int64_t i) { return I.sd + (m_ + 1) * m_ + (m_ + 1) + i; }
    double* MISC(Inst& I, int s) { return I.sd + (m_ + 1) * m_ + (m_ + 1) + m_ + s; }

    void setup(Inst& I, int64_t n, double lambda) {
        const size_t idx = (size_t)(&I - inst_.data());
        I.n = n;
        I.N = n * n * n;
        I.lambda = lambda;
        I.h2inv = (n + 1.0) * (n + 1.0);
        I.sd = sd_ + idx * (size_t)nsc_;
        I.ctl = ctl_ + 2 * idx;
        // hierarchy (poisson_mg.hpp:44-63 with even coarsening, as MG<T>)
        const numa_mg_config& g = cfg_.mg;
        int64_t mm = n;
        double h2 = I.h2inv;
        while (true) {
            Level L;
            L.n = mm;
            L.N = mm * mm * mm;
            L.h2inv = h2;
            I.lv.push_back(L);
            if (mm <= 1) break;
            if (!g.even_coarsening && (mm % 2 == 0)) break;
            if (mm <= std::max<int64_t>(1, g.min_coarse_n)) break;
            if (g.max_levels > 0 && (int64_t)I.lv.size() >= g.max_levels) break;
            mm = mm / 2;
            h2 /= 4.0;
        }
        require(I.lv.back().n <= 16,
                "the exact coarse solve needs a coarsest level of at most 16^3");
        // one slab per instance, every buffer 256-byte aligned
        size_t off = 0;
        std::vector<std::pair<void**, size_t>> parts;
        auto want = [&](void** dst, size_t bytes) { parts.push_back({dst, bytes}); };
        const size_t N = (size_t)I.N;
        want((void**)&I.xa, N * 8);
        want((void**)&I.x2, N * 8);
        want((void**)&I.Fx, N * 8);
        want((void**)&I.Fn, N * 8);
        want((void**)&I.z, N * 8);
        want((void**)&I.E, N * 4);
        I.V.assign((size_t)m_ + 1, nullptr);
        I.Z.assign((size_t)m_, nullptr);
        for (auto& v : I.V) want((void**)&v, N * 4);
        for (auto& v : I.Z) want((void**)&v, N * 4);
        want((void**)&I.partials, (size_t)kBatchPartials * 8);
        want((void**)&I.dst_tab, (size_t)kDstTab * 8);
        for (size_t l = 0; l < I.lv.size(); ++l) {
            Level& L = I.lv[l];
            const size_t NL = (size_t)L.N;
            want((void**)&L.x, NL * 4);
            want((void**)&L.x2, NL * 4);
            if (l > 0) want((void**)&L.b, NL * 4);
            want((void**)&L.d, NL * 4);
            if (cfg_.mg.smoother == NUMA_SMOOTH_CHEBYSHEV) {
                const double hi = cheb_lmax(L.n), lo = hi / cfg_.mg.cheb_ratio;
                cheb_step_coefs(L.h2inv, (int)cfg_.mg.degree, lo, hi, L.c1, L.c2);
            }
            if (l + 1 < I.lv.size()) L.rr_ich = resid_restrict_chunk<float>(L.n / 2, c_.cu_count);
        }
        for (auto& pr : parts) off += (pr.second + 255) / 256 * 256;
        void* slab = nullptr;
        NUMA_HIP_CHECK(hipMalloc(&slab, off));
        allocs_.push_back(slab);
        char* cur = static_cast<char*>(slab);
        for (auto& pr : parts) {
            *pr.first = cur;
            cur += (pr.second + 255) / 256 * 256;
        }
        const std::vector<double> tab = dst_table(I.lv.back().n);
        NUMA_HIP_CHECK(hipMemcpyAsync(I.dst_tab, tab.data(), tab.size() * sizeof(double),
                                      hipMemcpyHostToDevice, c_.stream));
        c_.sync();  // the table's host copy goes out of scope
    }

    // ---- device flags, reads, uploads ---------------------------------------
    void set_active(const std::vector<char>& on) {
        bool same = true;
        for (size_t i = 0; i < B_; ++i) same = same && act_[i] == (on[i] ? 1 : 0);
        if (same) return;
        sync();  // the pinned staging buffer is free again
        for (size_t i = 0; i < B_; ++i) {
            act_[i] = on[i] ? 1 : 0;
            hact_[i] = act_[i];
        }
        NUMA_HIP_CHECK(hipMemcpyAsync(act_d_, hact_, B_ * sizeof(int), hipMemcpyHostToDevice,
                                      c_.stream));
    }
    void sync() {
        c_.sync();
        ++syncs_;
    }
    // scalar slot at `slot_of(inst 0)`, `count` consecutive doubles, of every
    // instance -> hs_[i * count + k]
    void read_rows(const double* first, int64_t count) {
        NUMA_HIP_CHECK(hipMemcpy2DAsync(hs_, (size_t)count * sizeof(double), first,
                                        (size_t)nsc_ * sizeof(double),
                                        (size_t)count * sizeof(double), B_,
                                        hipMemcpyDeviceToHost, c_.stream));
        sync();
    }
    // hup_[i * count + k] -> `count` slots from `first` of every instance
    void write_rows(double* first, int64_t count) {
        NUMA_HIP_CHECK(hipMemcpy2DAsync(first, (size_t)nsc_ * sizeof(double), hup_,
                                        (size_t)count * sizeof(double),
                                        (size_t)count * sizeof(double), B_,
                                        hipMemcpyHostToDevice, c_.stream));
    }

    // ---- launches over the instance table -----------------------------------
    // fill(I, args) sets instance I's arguments and returns its block count;
    // bytes(I): the algorithmic HBM bytes of instance I (counted for the
    // active instances).
    template <class Op, class Fill, class Bytes>
    void launch(Fill&& fill, Bytes&& bytes) {
        using A = typename Op::Args;
        std::vector<A> tab(B_);
        BMap map{};
        map.count = (int)B_;
        unsigned tot = 0;
        for (size_t i = 0; i < B_; ++i) {
            map.start[i] = tot;
            tot += fill(inst_[i], tab[i]);
            const double b = bytes(inst_[i]);
            if (in_step_) step_bytes_[i] += b;  // per instance, counted when the step runs
            else if (!dry_ && act_[i] == 1) bytes_ += b;
        }
        map.start[B_] = tot;
        const A* dt = static_cast<const A*>(arena_.put(tab.data(), sizeof(A) * B_));
        if (dry_) return;
        hipLaunchKernelGGL((k_batch<Op>), dim3(tot), dim3(Op::kThreads), 0, c_.stream, dt, act_d_,
                           map);
        NUMA_HIP_LAUNCH_CHECK();
        ++launches_;
    }
    template <class Op, class Fill>
    void launch(Fill&& fill) {
        launch<Op>(fill, [](const Inst&) { return 0.0; });
    }
    // sum of instance I's `count(I)` partials and the epilogue `ep(I)`
    template <class Count, class Ep>
    void finish(Count&& count, Ep&& ep) {
        launch<OpFinish>([&](Inst& I, OpFinish::Args& a) {
            a.partials = I.partials;
            a.count = count(I);
            require(a.count >= 1 && a.count <= kBatchPartials, "bad batched reduction size");
            a.ep = ep(I);
            return 1u;
        });
    }
    static unsigned stencil_tiles(const Inst& I, int64_t n) {
        const StencilShape s = stencil_shape(n, false);
        (void)I;
        return (unsigned)(s.gx * s.gy * s.gz);
    }
    // resident workgroups of the LDS-tiled form on the device (its plane chunk)
    template <bool REDUCE, class Load, class Epi>
    int64_t lds_slots() const {
        static const int per_cu =
            blocks_per_cu(k_stencil_lds<kLdsRows, REDUCE, true, Load, Epi>, 256);
        return (int64_t)per_cu * c_.cu_count;
    }
    // the stencil launch of each instance's level-`lvl` grid (launch_stencil's
    // form for its size); make(I) -> {loader, epilogue}
    template <bool REDUCE, class Load, class Epi, class Make, class Bytes>
    void stencil(int lvl, Make&& make, Bytes&& bytes) {
        launch<OpStencil<REDUCE, Load, Epi>>(
            [&](Inst& I, typename OpStencil<REDUCE, Load, Epi>::Args& a) {
                const int64_t n = lvl < 0 ? I.n : I.lv[(size_t)lvl].n;
                const bool lds = n >= kPlanesMinN;
                const StencilShape s =
                    lds ? stencil_shape_lds(n, lds_slots<REDUCE, Load, Epi>())
                        : stencil_shape(n, Epi::kMarch);
                a.n = (int32_t)n;
                a.ichunk = s.ichunk;
                a.gx = s.gx;
                a.gy = s.gy;
                a.gz = s.gz;
                a.tk = lds ? 0 : s.tk;
                const auto le = make(I);
                a.ldr = le.first;
                a.epi = le.second;
                a.red = Reduce{I.partials};
                return s.blocks;
            },
            bytes);
    }
    template <class TV, class Vec, class Slot>
    void norm(Vec&& vec, Slot&& slot, bool scale_slot0 = false) {
        launch<OpDot<TV, TV>>(
            [&](Inst& I, typename OpDot<TV, TV>::Args& a) {
                a.N = I.N;
                a.a = vec(I);
                a.b = a.a;
                a.red = Reduce{I.partials};
                return stream_blocks(I.N);
            },
            [](const Inst& I) { return (double)I.N * sizeof(TV); });
        finish([](Inst& I) { return (int64_t)stream_blocks(I.N); },
               [&](Inst& I) {
                   return scale_slot0 ? Epilogue{EPI_NORM_SCALE, 1.0, nullptr, slot(I), S(I, 0)}
                                      : Epilogue{EPI_SQRT, 1.0, nullptr, slot(I), nullptr};
               });
    }
    // y = v over each instance's level-`lvl` grid (-1: the finest)
    template <class T>
    void fill(double v, std::function<T*(Inst&)> y, int lvl = -1) {
        auto len = [lvl](const Inst& I) { return lvl < 0 ? I.N : I.lv[(size_t)lvl].N; };
        launch<OpFill<T>>(
            [&](Inst& I, typename OpFill<T>::Args& a) {
                a.N = len(I);
                a.y = y(I);
                a.v = v;
                return stream_blocks(a.N);
            },
            [&](const Inst& I) { return (double)len(I) * sizeof(T); });
    }
    void copy_x(std::function<const double*(Inst&)> src, std::function<double*(Inst&)> dst) {
        launch<OpConvert<double, double>>(
            [&](Inst& I, OpConvert<double, double>::Args& a) {
                a.N = I.N;
                a.x = src(I);
                a.y = dst(I);
                return stream_blocks(I.N);
            },
            [](const Inst& I) { return 16.0 * (double)I.N; });
    }

    // F at x_c (trial = false; F -> F_c) or at the trial point x_c - a z with
    // a = the S_LSA slot (F -> F_t, the point -> x_t); ||F|| -> S_FNORM
    // (eval_F of the single-instance solver).
    void eval_F(bool trial) {
        const auto ep = [&](Inst& I) { return Epilogue{EPI_SQRT, 1.0, nullptr, MISC(I, S_FNORM), nullptr}; };
        if (!trial) {
            stencil<true, LdVec<double>, EpF<false>>(
                -1,
                [&](Inst& I) {
                    return std::make_pair(LdVec<double>{I.xc},
                                          EpF<false>{I.h2inv, I.lambda, I.Fc, nullptr,
                                                     MISC(I, S_LAMBDA)});
                },
                [](const Inst& I) { return 16.0 * (double)I.N; });
        } else {
            stencil<true, LdStepDev, EpF<true>>(
                -1,
                [&](Inst& I) {
                    return std::make_pair(LdStepDev{I.xc, I.z, MISC(I, S_LSA), 0.0},
                                          EpF<true>{I.h2inv, I.lambda, I.Ft, I.xt,
                                                    MISC(I, S_LAMBDA)});
                },
                [](const Inst& I) { return 32.0 * (double)I.N; });
        }
        finish([&](Inst& I) { return (int64_t)stencil_tiles(I, I.n); }, ep);
    }

    // ||x_c|| and lambda e^{x_c} in FP32 (point_setup)
    void point_setup() {
        launch<OpXnormExp>(
            [&](Inst& I, OpXnormExp::Args& a) {
                a.N = I.N;
                a.x = I.xc;
                a.lambda = I.lambda;
                a.lam_ptr = MISC(I, S_LAMBDA);
                a.E = I.E;
                a.red = Reduce{I.partials};
                return stream_blocks(I.N);
            },
            [](const Inst& I) { return 12.0 * (double)I.N; });
        finish([](Inst& I) { return (int64_t)stream_blocks(I.N); },
               [&](Inst& I) { return Epilogue{EPI_SQRT, 1.0, nullptr, MISC(I, S_XNORM), nullptr}; });
    }

    // ---- V-cycle (MG<float>::apply_norm with one launch per smoothing step) -
    bool final_done_ = false;
    void poly_step(size_t l, int k, bool wj, bool use_d, bool write_d, bool from_zero, bool dst) {
        auto c1 = [&](Inst& I) { return wj ? 0.0 : I.lv[l].c1[(size_t)k]; };
        auto c2 = [&](Inst& I) {
            return wj ? cfg_.mg.omega / (6.0 * I.lv[l].h2inv) : I.lv[l].c2[(size_t)k];
        };
        auto bof = [&](Inst& I) -> const float* { return l == 0 ? I.vb : I.lv[l].b; };
        auto xo = [&](Inst& I) { return dst ? I.vout : I.lv[l].x2; };
        if (from_zero) {
            launch<OpPolyFirst<float>>(
                [&](Inst& I, OpPolyFirst<float>::Args& a) {
                    a.N = I.lv[l].N;
                    a.c2 = c2(I);
                    a.b = bof(I);
                    a.d = write_d ? I.lv[l].d : nullptr;
                    a.xn = xo(I);
                    return stream_blocks(I.lv[l].N);
                },
                [&](const Inst& I) { return (write_d ? 12.0 : 8.0) * (double)I.lv[l].N; });
        } else {
            auto run = [&](auto ud, auto wd) {
                constexpr bool UD = decltype(ud)::value, WD = decltype(wd)::value;
                stencil<false, LdVec<float>, EpPoly<float, UD, WD>>(
                    (int)l,
                    [&](Inst& I) {
                        return std::make_pair(LdVec<float>{I.lv[l].x},
                                              EpPoly<float, UD, WD>{I.lv[l].h2inv, c1(I), c2(I),
                                                                    bof(I), I.lv[l].d, xo(I)});
                    },
                    [&](const Inst& I) {
                        return (12.0 + (UD ? 4.0 : 0.0) + (WD ? 4.0 : 0.0)) * (double)I.lv[l].N;
                    });
            };
            using TT = std::true_type;
            using FF = std::false_type;
            if (use_d && write_d) run(TT{}, TT{});
            else if (use_d) run(TT{}, FF{});
            else if (write_d) run(FF{}, TT{});
            else run(FF{}, FF{});
        }
        if (dst) {
            final_done_ = true;
        } else {
            for (auto& I : inst_) std::swap(I.lv[l].x, I.lv[l].x2);
        }
    }
    // One sweep (the degree's steps) per launch in the box form of shape
    // `shape` (bit for bit the per-step launches).
    template <int DEG, int SH>
    void box_sweep(size_t l, bool from_zero, bool dst) {
        launch<OpPolyBox<float, DEG, SH>>(
            [&](Inst& I, typename OpPolyBox<float, DEG, SH>::Args& a) {
                const Level& L = I.lv[l];
                int tk = 0, tj = 0, total = 0;
                const dim3 g = box_grid<OpPolyBox<float, DEG, SH>::EI, OpPolyBox<float, DEG, SH>::EJ,
                                        OpPolyBox<float, DEG, SH>::EK>((int)L.n, tk, tj, total);
                a.n = (int32_t)L.n;
                a.tiles_k = tk;
                a.tiles_j = tj;
                a.total = total;
                a.h2inv = L.h2inv;
                for (int q = 0; q < DEG; ++q) {
                    a.cs.c1[q] = q == 0 ? 0.0 : L.c1[(size_t)q];
                    a.cs.c2[q] = L.c2[(size_t)q];
                }
                a.x = from_zero ? nullptr : L.x;
                a.b = l == 0 ? I.vb : L.b;
                a.xo = dst ? I.vout : L.x2;
                return g.x;
            },
            [&](const Inst& I) { return (from_zero ? 8.0 : 12.0) * (double)I.lv[l].N; });
    }
    // One sweep per launch in the plane-marching form (R rows per thread,
    // `ichunk` planes per workgroup).
    template <int DEG, int R>
    void march_sweep(size_t l, int ichunk, bool from_zero, bool dst) {
        launch<OpPolyFused<float, DEG, R>>(
            [&](Inst& I, typename OpPolyFused<float, DEG, R>::Args& a) {
                const Level& L = I.lv[l];
                int tk = 0, tj = 0, ch = 0;
                const int ic = (int)std::min<int64_t>(ichunk, L.n);
                const dim3 g = fused_grid<DEG, R>((int)L.n, ic, tk, tj, ch);
                a.n = (int32_t)L.n;
                a.ichunk = ic;
                a.tiles_k = tk;
                a.tiles_j = tj;
                a.chunks = ch;
                a.h2inv = L.h2inv;
                for (int q = 0; q < DEG; ++q) {
                    a.cs.c1[q] = q == 0 ? 0.0 : L.c1[(size_t)q];
                    a.cs.c2[q] = L.c2[(size_t)q];
                }
                a.x = from_zero ? nullptr : L.x;
                a.b = l == 0 ? I.vb : L.b;
                a.xo = dst ? I.vout : L.x2;
                return g.x;
            },
            [&](const Inst& I) { return (from_zero ? 8.0 : 12.0) * (double)I.lv[l].N; });
    }
    template <int DEG>
    void sweep_form(size_t l, int form, bool from_zero, bool dst) {
        switch (form) {
            case 1: box_sweep<DEG, 1>(l, from_zero, dst); break;
            case 2: box_sweep<DEG, 2>(l, from_zero, dst); break;
            case 3: box_sweep<DEG, 3>(l, from_zero, dst); break;
            case 4: march_sweep<DEG, 2>(l, 16, from_zero, dst); break;
            case 5: march_sweep<DEG, 2>(l, 32, from_zero, dst); break;
            case 6: march_sweep<DEG, 4>(l, 16, from_zero, dst); break;
            default: march_sweep<DEG, 4>(l, 32, from_zero, dst); break;
        }
    }
    template <int DEG>
    void box_sweep_deg(size_t l, int shape, bool from_zero, bool dst) {
        switch (shape) {
            case 1: box_sweep<DEG, 1>(l, from_zero, dst); break;
            case 2: box_sweep<DEG, 2>(l, from_zero, dst); break;
            default: box_sweep<DEG, 3>(l, from_zero, dst); break;
        }
    }
    bool box_ok() const {
        return cfg_.mg.smoother == NUMA_SMOOTH_CHEBYSHEV && cfg_.mg.degree >= 1 &&
               cfg_.mg.degree <= 3;
    }
    void smooth(size_t l, int sweeps, bool from_zero, bool last_dst) {
        if (sweeps == 0) {
            if (from_zero) fill<float>(0.0, [&](Inst& I) { return I.lv[l].x; }, (int)l);
            return;
        }
        if (box_ok() && lvl_box_[l] > 0) {
            for (int s = 0; s < sweeps; ++s) {
                const bool dst = last_dst && s + 1 == sweeps;
                const bool fz = from_zero && s == 0;
                switch (cfg_.mg.degree) {
                    case 1: sweep_form<1>(l, lvl_box_[l], fz, dst); break;
                    case 2: sweep_form<2>(l, lvl_box_[l], fz, dst); break;
                    default: sweep_form<3>(l, lvl_box_[l], fz, dst); break;
                }
                if (dst) final_done_ = true;
                else
                    for (auto& I : inst_) std::swap(I.lv[l].x, I.lv[l].x2);
            }
            return;
        }
        const bool wj = cfg_.mg.smoother == NUMA_SMOOTH_WJACOBI;
        const int deg = wj ? 1 : (int)cfg_.mg.degree;
        for (int s = 0; s < sweeps; ++s)
            for (int k = 0; k < deg; ++k) {
                const bool dst = last_dst && s + 1 == sweeps && k + 1 == deg;
                if (wj)
                    poly_step(l, 0, true, false, false, from_zero && s == 0, dst);
                else
                    poly_step(l, k, false, k > 0, k + 1 < deg, from_zero && s == 0 && k == 0, dst);
            }
    }
    void cycle(size_t l) {
        const size_t last = inst_[0].lv.size() - 1;
        if (l == last) {
            launch<OpCoarseDst<float>>(
                [&](Inst& I, OpCoarseDst<float>::Args& a) {
                    const Level& L = I.lv[l];
                    a.n = L.n;
                    a.h2inv = L.h2inv;
                    a.tab = I.dst_tab;
                    a.b = l == 0 ? I.vb : L.b;
                    a.x = L.x;
                    return 1u;
                },
                [&](const Inst& I) { return 8.0 * (double)I.lv[l].N; });
            return;
        }
        smooth(l, (int)cfg_.mg.pre, true, false);
        resid_restrict(l);
        cycle(l + 1);
        prolong(l);
        smooth(l, (int)cfg_.mg.post, false, l == 0);
    }
    // b_{l+1} = R (b_l - A x_l) in one pass (MG<T>::resid_restrict)
    void resid_restrict(size_t l) {
        launch<OpResidRestrict<float>>(
            [&](Inst& I, OpResidRestrict<float>::Args& a) {
                const Level& L = I.lv[l];
                const Level& C = I.lv[l + 1];
                a.nf = (int32_t)L.n;
                a.nc = (int32_t)C.n;
                a.ich = lvl_rr_[l] > 0 ? (int32_t)std::min<int64_t>(lvl_rr_[l], C.n) : L.rr_ich;
                a.tiles_k = (int32_t)((C.n + kRRK - 1) / kRRK);
                a.tiles_j = (int32_t)((C.n + kRRJ - 1) / kRRJ);
                a.chunks = (int32_t)((C.n + a.ich - 1) / a.ich);
                a.h2inv = L.h2inv;
                a.x = L.x;
                a.b = l == 0 ? I.vb : L.b;
                a.bc = C.b;
                const int64_t total = (int64_t)a.tiles_k * a.tiles_j * a.chunks;
                return (unsigned)((total + 7) / 8 * 8);
            },
            [&](const Inst& I) { return 8.5 * (double)I.lv[l].N; });
    }
    // x_l += P x_{l+1}
    void prolong(size_t l) {
        launch<OpProlong<float>>(
            [&](Inst& I, OpProlong<float>::Args& a) {
                const Level& L = I.lv[l];
                const Level& C = I.lv[l + 1];
                const int64_t ncell = (L.n + 1) / 2;
                a.nf = (int32_t)L.n;
                a.nc = (int32_t)C.n;
                a.gx = (int32_t)((ncell + 63) / 64);
                a.gy = (int32_t)((ncell + 3) / 4);
                a.xc = C.x;
                a.xf = L.x;
                return (unsigned)((int64_t)a.gx * a.gy * ncell);
            },
            [&](const Inst& I) { return 8.5 * (double)I.lv[l].N; });
    }
    // Smoothing form per level (all instances): time one pre-smoothing phase
    // from zero (b = 1) for one launch per step and each box shape, keep the
    // fastest (a later candidate must win by 2%), remember it in the tune
    // cache under the level's grid sizes.
    void autotune() {
        tuned_ = true;
        const size_t nl = inst_[0].lv.size();
        lvl_box_.assign(nl, 0);
        lvl_rr_.assign(nl, 0);
        if (!box_ok() || cfg_.mg.fuse == NUMA_FUSE_NONE || cfg_.mg.pre == 0) return;
        const std::string dev = TuneCache::device_key(c_.device);
        const std::vector<char> all(B_, 1);
        set_active(all);
        for (size_t l = 0; l + 1 < nl; ++l) {
            std::map<int64_t, int> sizes;
            for (auto& I : inst_) ++sizes[I.lv[l].n];
            std::ostringstream key;
            key << dev << "|batch-smooth-v4|f32|deg=" << cfg_.mg.degree << "|pre=" << cfg_.mg.pre
                << "|n=";
            for (auto& kv : sizes) key << kv.first << "x" << kv.second << ",";
            std::string val;
            const auto held = TuneCache::get().tuning(key.str());
            if (TuneCache::get().lookup(key.str(), val)) {
                int v = -1, r = -1;
                if (std::sscanf(val.c_str(), "%d,%d", &v, &r) == 2 && v >= 0 && v <= kSweepForms &&
                    r >= 0 && r <= 1024) {
                    lvl_box_[l] = v;
                    lvl_rr_[l] = r;
                    continue;
                }
            }
            // b = 1 on the level (level 0: V_0 stands in for the input)
            for (auto& I : inst_) {
                I.vb = I.V[0];
                I.vout = I.Z[0];
            }
            if (l == 0)
                fill<float>(1.0, [](Inst& I) { return I.V[0]; });
            else
                fill<float>(1.0, [&](Inst& I) { return I.lv[l].b; }, (int)l);
            hipEvent_t e0, e1;
            NUMA_HIP_CHECK(hipEventCreate(&e0));
            NUMA_HIP_CHECK(hipEventCreate(&e1));
            const bool log = std::getenv("NUMA_HIP_TUNE_LOG") != nullptr;
            double best = 1e300;
            int pick = 0;
            for (int cand = 0; cand <= kSweepForms; ++cand) {
                lvl_box_[l] = cand;
                smooth(l, (int)cfg_.mg.pre, true, false);  // warm-up
                float ms = 1e30f;
                for (int trial = 0; trial < 5; ++trial) {
                    NUMA_HIP_CHECK(hipEventRecord(e0, c_.stream));
                    for (int rep = 0; rep < 3; ++rep) smooth(l, (int)cfg_.mg.pre, true, false);
                    NUMA_HIP_CHECK(hipEventRecord(e1, c_.stream));
                    NUMA_HIP_CHECK(hipEventSynchronize(e1));
                    float t = 0.0f;
                    NUMA_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
                    ms = std::min(ms, t);
                }
                if (log)
                    std::fprintf(stderr, "numa_hip batch tune: level %zu form %d %.2f us\n", l,
                                 cand, ms * 1e3 / 3.0);
                if (ms < 0.98 * best) {
                    best = ms;
                    pick = cand;
                }
            }
            lvl_box_[l] = pick;
            // coarse planes per workgroup of the residual / restriction
            // (every chunk gives the same values): the lone instance's choice
            // (0) or a fixed chunk
            double best_rr = 1e300;
            int pick_rr = 0;
            for (int ich : {0, 2, 4, 8, 16, 32}) {
                lvl_rr_[l] = ich;
                resid_restrict(l);  // warm-up
                float ms = 1e30f;
                for (int trial = 0; trial < 5; ++trial) {
                    NUMA_HIP_CHECK(hipEventRecord(e0, c_.stream));
                    for (int rep = 0; rep < 3; ++rep) resid_restrict(l);
                    NUMA_HIP_CHECK(hipEventRecord(e1, c_.stream));
                    NUMA_HIP_CHECK(hipEventSynchronize(e1));
                    float t = 0.0f;
                    NUMA_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
                    ms = std::min(ms, t);
                }
                if (log)
                    std::fprintf(stderr, "numa_hip batch tune: level %zu resid/restrict chunk %d %.2f us\n",
                                 l, ich, ms * 1e3 / 3.0);
                if (ms < 0.98 * best_rr) {
                    best_rr = ms;
                    pick_rr = ich;
                }
            }
            lvl_rr_[l] = pick_rr;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            TuneCache::get().store(key.str(), std::to_string(pick) + "," + std::to_string(pick_rr));
        }
        bytes_ = 0.0;
        launches_ = 0;
    }

    // z_j = V-cycle(V_j), ||z_j|| -> S_VNORM
    void vcycle(size_t j) {
        for (auto& I : inst_) {
            I.vb = I.V[j];
            I.vout = I.Z[j];
        }
        final_done_ = false;
        cycle(0);
        if (!final_done_)
            launch<OpConvert<float, float>>(
                [&](Inst& I, OpConvert<float, float>::Args& a) {
                    a.N = I.N;
                    a.x = I.lv[0].x;
                    a.y = I.vout;
                    return stream_blocks(I.N);
                },
                [](const Inst& I) { return 8.0 * (double)I.N; });
        norm<float>([&](Inst& I) { return (const float*)I.vout; },
                    [&](Inst& I) { return MISC(I, S_VNORM); });
    }

    // ---- one Arnoldi step j of every active instance ------------------------
    void arnoldi_body(size_t j) {
        vcycle(j);
        const double root = std::sqrt(std::numeric_limits<double>::epsilon());
        stencil<true, LdVec<float>, EpJvFDe<float, true, false>>(
            -1,
            [&](Inst& I) {
                const FDStep st{root, MISC(I, S_XNORM), MISC(I, S_VNORM)};
                return std::make_pair(LdVec<float>{I.Z[j]},
                                      EpJvFDe<float, true, false>{I.h2inv, st, S(I, j), I.E,
                                                                  I.V[j + 1], I.V[0], nullptr, 0.0,
                                                                  1.0});
            },
            [](const Inst& I) { return 16.0 * (double)I.N; });
        finish([&](Inst& I) { return (int64_t)stencil_tiles(I, I.n); },
               [&](Inst& I) { return Epilogue{EPI_STORE, 1.0, S(I, 0), H(I, 0, j), nullptr}; });
        for (size_t i = 0; i <= j; ++i) {
            const bool last = i == j;
            auto args = [&](Inst& I, auto& a) {
                a.N = I.N;
                a.w = I.V[j + 1];
                a.Vi = I.V[i];
                a.hi = H(I, i, j);
                a.si = S(I, i);
                a.Vnext = last ? nullptr : I.V[i + 1];
                a.red = Reduce{I.partials};
                return stream_blocks(I.N);
            };
            if (last)
                launch<OpMgs<true, float>>(args, [](const Inst& I) { return 12.0 * (double)I.N; });
            else
                launch<OpMgs<false, float>>(args, [](const Inst& I) { return 16.0 * (double)I.N; });
            finish([](Inst& I) { return (int64_t)stream_blocks(I.N); },
                   [&](Inst& I) {
                       return last ? Epilogue{EPI_NORM_SCALE, 1.0, nullptr, H(I, j + 1, j), S(I, j + 1)}
                                   : Epilogue{EPI_STORE, 1.0, S(I, i + 1), H(I, i + 1, j), nullptr};
                   });
        }
    }
    // Runs step j for the active instances: replays its graph (captured at
    // first use: a dry pass uploads the argument tables, then the body is
    // captured), or launches eagerly.
    void run_step(size_t j) {
        ++steps_;
        std::fill(step_bytes_.begin(), step_bytes_.end(), 0.0);
        struct InStep {
            bool& f;
            explicit InStep(bool& g) : f(g) { f = true; }
            ~InStep() { f = false; }
        } guard(in_step_);
        if (!graphs_) {
            arnoldi_body(j);
            count_bytes(step_bytes_);
            return;
        }
        auto it = gcache_.find(j);
        if (it == gcache_.end()) {
            // the dry pass swaps the level buffers as the real one will: put
            // them back so that the captured body finds its tables cached
            std::vector<std::pair<float*, float*>> saved;
            for (auto& I : inst_)
                for (auto& L : I.lv) saved.push_back({L.x, L.x2});
            dry_ = true;
            try {
                arnoldi_body(j);
            } catch (...) {
                dry_ = false;
                throw;
            }
            dry_ = false;
            size_t q = 0;
            for (auto& I : inst_)
                for (auto& L : I.lv) {
                    L.x = saved[q].first;
                    L.x2 = saved[q].second;
                    ++q;
                }
            sync();
            std::fill(step_bytes_.begin(), step_bytes_.end(), 0.0);
            const int64_t l0 = launches_;
            NUMA_HIP_CHECK(hipStreamBeginCapture(c_.stream, hipStreamCaptureModeThreadLocal));
            try {
                arnoldi_body(j);
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
            launches_ = l0;  // captured launches run inside the graph
            it = gcache_.emplace(j, Captured{ge, step_bytes_}).first;
        }
        NUMA_HIP_CHECK(hipGraphLaunch(it->second.exec, c_.stream));
        ++graph_launches_;
        count_bytes(it->second.bytes);
    }
    void count_bytes(const std::vector<double>& per_inst) {
        for (size_t i = 0; i < B_; ++i)
            if (act_[i] == 1) bytes_ += per_inst[i];
    }
    void drop_graphs() {
        for (auto& kv : gcache_) (void)hipGraphExecDestroy(kv.second.exec);
        gcache_.clear();
    }

    void gmres_batch(const std::vector<char>& R, const std::vector<double>& eta, int64_t restart,
                     int64_t max_iters);
    size_t idx(const Inst& I) const { return (size_t)(&I - inst_.data()); }

    Ctx& c_;
    numa_solver_config cfg_;
    int64_t m_;
    size_t B_ = 0;
    int64_t nsc_ = 0;
    std::vector<Inst> inst_;
    std::vector<GState> gst_;
    double* sd_ = nullptr;  // scalar slots, nsc_ per instance
    int* ctl_ = nullptr;    // combine control, 2 per instance
    int* act_d_ = nullptr;  // device flags
    double* hs_ = nullptr;  // pinned: reads
    double* hup_ = nullptr; // pinned: uploads
    int* hctl_ = nullptr;
    int* hact_ = nullptr;
    std::vector<int> act_;  // flags as last uploaded (-1: unknown)
    ArgArena arena_;
    bool dry_ = false;
    bool in_step_ = false;
    bool tuned_ = false;
    std::vector<int> lvl_box_;  // per level: smoothing form (0 one launch per step, set_smoothing)
    std::vector<int> lvl_rr_;   // per level: residual / restriction chunk (0: per instance)
    bool graphs_ = false;
    struct Captured {
        hipGraphExec_t exec;
        std::vector<double> bytes;  // per instance, algorithmic
    };
    std::map<size_t, Captured> gcache_;
    std::vector<double> step_bytes_ = std::vector<double>(kBatchMax, 0.0);
    std::vector<void*> allocs_;
    int64_t launches_ = 0, graph_launches_ = 0, steps_ = 0, syncs_ = 0;
    double bytes_ = 0.0;
};

// The GMRES solves of the instances in R (rtol eta[i]) at their Newton
// points: J(x_c) z = F(x_c) from z = 0 into z (linalg/gmres.hpp:61-223 as in
// BratuSolver::gmres with FGMRES, the V-cycle and a zero guess).
void BatchSolver::gmres_batch(const std::vector<char>& R, const std::vector<double>& eta,
                              int64_t restart, int64_t max_iters) {
    const double root = std::sqrt(std::numeric_limits<double>::epsilon());
    const size_t m = (size_t)restart;
    std::vector<char> on(B_, 0);
    for (size_t i = 0; i < B_; ++i) {
        GState& G = gst_[i];
        G.res = numa_gmres_result{};
        G.in = R[i] != 0;
        G.first = true;
        G.xl_unset = true;
        G.cyc = false;
        G.rtol = eta[i];
    }
    // From a zero guess the first residual is b = F(x_c): ||b||, 1/||b|| and
    // V_0 = b in one pass.
    set_active(R);
    launch<OpConvertNorm<double, float>>(
        [&](Inst& I, OpConvertNorm<double, float>::Args& a) {
            a.N = I.N;
            a.x = I.Fc;
            a.y = I.V[0];
            a.red = Reduce{I.partials};
            return stream_blocks(I.N);
        },
        [](const Inst& I) { return 12.0 * (double)I.N; });
    finish([](Inst& I) { return (int64_t)stream_blocks(I.N); },
           [&](Inst& I) { return Epilogue{EPI_NORM_SCALE, 1.0, nullptr, MISC(I, S_BNORM), S(I, 0)}; });
    read_rows(MISC(inst_[0], S_BNORM), 1);
    std::vector<char> zero(B_, 0);
    bool any_zero = false;
    for (size_t i = 0; i < B_; ++i) {
        if (!R[i]) continue;
        gst_[i].bnorm = hs_[i];
        if (gst_[i].bnorm == 0.0) {
            zero[i] = 1;
            any_zero = true;
            gst_[i].res.converged = 1;
            gst_[i].in = false;
        }
    }
    if (any_zero) {
        set_active(zero);
        fill<double>(0.0, [](Inst& I) { return I.z; });
    }
    auto any = [&](auto pred) {
        for (size_t i = 0; i < B_; ++i)
            if (pred(i)) return true;
        return false;
    };
    while (any([&](size_t i) { return gst_[i].in; })) {
        // cycle start: beta = ||b - J x_l|| (b for the first cycle)
        std::vector<char> need(B_, 0);
        bool any_need = false;
        for (size_t i = 0; i < B_; ++i) {
            GState& G = gst_[i];
            if (!G.in) continue;
            if (G.first) {
                G.beta = G.bnorm;
                G.first = false;
            } else {
                need[i] = 1;
                any_need = true;
            }
        }
        if (any_need) {
            // x_l is set after every cycle that ran (a cycle runs >= 1 step)
            set_active(need);
            norm<double>([](Inst& I) { return (const double*)I.z; },
                         [&](Inst& I) { return MISC(I, S_VNORM); });
            stencil<true, LdVec<double>, EpJvFDe<float, false, true>>(
                -1,
                [&](Inst& I) {
                    const FDStep st{root, MISC(I, S_XNORM), MISC(I, S_VNORM)};
                    return std::make_pair(LdVec<double>{I.z},
                                          EpJvFDe<float, false, true>{I.h2inv, st, nullptr, I.E,
                                                                      I.V[0], nullptr, I.Fc, 0.0,
                                                                      1.0});
                },
                [](const Inst& I) { return 24.0 * (double)I.N; });
            finish([&](Inst& I) { return (int64_t)stencil_tiles(I, I.n); },
                   [&](Inst& I) {
                       return Epilogue{EPI_NORM_SCALE, 1.0, nullptr, MISC(I, S_BETA), S(I, 0)};
                   });
            read_rows(MISC(inst_[0], S_BETA), 1);
            for (size_t i = 0; i < B_; ++i)
                if (need[i]) {
                    gst_[i].beta = hs_[i];
                    ++gst_[i].res.matvecs;
                }
        }
        for (size_t i = 0; i < B_; ++i) {
            GState& G = gst_[i];
            if (!G.in) continue;
            G.res.rel_res = G.beta / G.bnorm;
            if (G.res.rel_res <= G.rtol) {
                G.res.converged = 1;
                G.in = false;
                continue;
            }
            for (auto& row : G.Hraw) std::fill(row.begin(), row.end(), 0.0);
            std::fill(G.g.begin(), G.g.end(), 0.0);
            G.g[0] = G.beta;
            G.s[0] = G.beta > 0.0 ? 1.0 / G.beta : 0.0;
            G.k = 0;
            G.cyc = true;
        }
        // lockstep Arnoldi steps of the cycle
        for (size_t j = 0; j < m; ++j) {
            for (size_t i = 0; i < B_; ++i) on[i] = gst_[i].cyc ? 1 : 0;
            if (!any([&](size_t i) { return on[i] != 0; })) break;
            set_active(on);
            run_step(j);
            NUMA_HIP_CHECK(hipMemcpy2DAsync(hs_, (j + 2) * sizeof(double), H(inst_[0], 0, j),
                                            (size_t)nsc_ * sizeof(double),
                                            (j + 2) * sizeof(double), B_, hipMemcpyDeviceToHost,
                                            c_.stream));
            sync();
            for (size_t i = 0; i < B_; ++i) {
                if (!on[i]) continue;
                GState& G = gst_[i];
                auto& Hh = G.Hh;
                ++G.res.precs;
                ++G.res.matvecs;
                for (size_t r = 0; r <= j + 1; ++r) Hh[r][j] = hs_[i * (j + 2) + r];
                G.s[j + 1] = Hh[j + 1][j] > 0.0 ? 1.0 / Hh[j + 1][j] : 0.0;
                for (size_t r = 0; r <= j + 1; ++r) G.Hraw[r][j] = Hh[r][j];
                for (size_t r = 0; r < j; ++r) {
                    const double t = G.cs[r] * Hh[r][j] + G.sn[r] * Hh[r + 1][j];
                    Hh[r + 1][j] = -G.sn[r] * Hh[r][j] + G.cs[r] * Hh[r + 1][j];
                    Hh[r][j] = t;
                }
                const double denom = std::hypot(Hh[j][j], Hh[j + 1][j]);
                if (!(denom > 0.0) || !std::isfinite(denom)) {
                    G.res.breakdown = 1;
                    G.in = false;
                    G.cyc = false;
                    continue;
                }
                G.cs[j] = Hh[j][j] / denom;
                G.sn[j] = Hh[j + 1][j] / denom;
                Hh[j][j] = denom;
                Hh[j + 1][j] = 0.0;
                G.g[j + 1] = -G.sn[j] * G.g[j];
                G.g[j] = G.cs[j] * G.g[j];
                ++G.k;
                ++G.res.iters;
                G.res.rel_res = std::abs(G.g[j + 1]) / G.bnorm;
                if (G.res.rel_res <= G.rtol || G.k == m || G.res.iters >= max_iters)
                    G.cyc = false;
            }
        }
        // end of the cycle: x_l (=|+=) sum y_i s_i Z_i for the instances that ran it
        std::vector<char> comb(B_, 0);
        bool any_comb = false;
        for (size_t i = 0; i < B_; ++i) {
            GState& G = gst_[i];
            if (!G.in || G.k == 0) continue;
            const size_t k = G.k;
            bool broke = false;
            for (size_t ii = k; ii-- > 0;) {
                const double pivot = G.Hh[ii][ii];
                if (!(std::abs(pivot) > 0.0) || !std::isfinite(pivot)) {
                    broke = true;
                    break;
                }
                double acc = G.g[ii];
                for (size_t jj = ii + 1; jj < k; ++jj) acc -= G.Hh[ii][jj] * G.y[jj];
                G.y[ii] = acc / pivot;
            }
            if (broke) {
                G.res.breakdown = 1;
                G.in = false;
                continue;
            }
            for (size_t q = 0; q < k; ++q) hup_[i * (size_t)m_ + q] = G.y[q] * G.s[q];
            for (size_t q = k; q < (size_t)m_; ++q) hup_[i * (size_t)m_ + q] = 0.0;
            hctl_[2 * i] = (int)k;
            hctl_[2 * i + 1] = G.xl_unset ? 0 : 1;  // 0 + sum: same bits as a fill first
            comb[i] = 1;
            any_comb = true;
        }
        if (any_comb) {
            write_rows(COEF(inst_[0], 0), m_);
            NUMA_HIP_CHECK(hipMemcpyAsync(ctl_, hctl_, B_ * 2 * sizeof(int),
                                          hipMemcpyHostToDevice, c_.stream));
            set_active(comb);  // synchronises before switching flags: the uploads are read
            launch<OpCombineDev<float, double>>(
                [&](Inst& I, OpCombineDev<float, double>::Args& a) {
                    a.N = I.N;
                    for (size_t q = 0; q < (size_t)m_; ++q) a.v[q] = I.Z[q];
                    a.coef = COEF(I, 0);
                    a.ctl = I.ctl;
                    a.y = I.z;
                    return stream_blocks(I.N);
                },
                [&](const Inst& I) {
                    const size_t i = idx(I);
                    return (4.0 * (double)gst_[i].k + 8.0 + (gst_[i].xl_unset ? 0.0 : 8.0)) *
                           (double)I.N;
                });
            sync();  // the pinned uploads are free again
            for (size_t i = 0; i < B_; ++i) {
                if (!comb[i]) continue;
                GState& G = gst_[i];
                G.xl_unset = false;
                if (G.res.rel_res <= G.rtol) {
                    G.res.converged = 1;
                    G.in = false;
                } else if (G.res.iters >= max_iters) {
                    G.in = false;
                }
            }
        }
    }
    // x_l = 0 where no cycle set it (settle_xl)
    std::vector<char> unset(B_, 0);
    bool any_unset = false;
    for (size_t i = 0; i < B_; ++i)
        if (R[i] && gst_[i].xl_unset && !zero[i]) {
            unset[i] = 1;
            any_unset = true;
        }
    if (any_unset) {
        set_active(unset);
        fill<double>(0.0, [](Inst& I) { return I.z; });
    }
}

// newton_krylov (methods/krylov_newton.hpp:69-114) for every instance, as
// BratuSolver::solve: Eisenstat-Walker forcing with the final-solve
// safeguard, FGMRES from zero, Armijo backtracking.
void BatchSolver::solve(const numa_krylov_params& p, double* const* x, numa_krylov_result* out) {
    require(p.gmres_restart >= 1 && p.gmres_restart <= m_,
            "batched solve: gmres_restart must be in [1, the batch's restart]");
    require(p.gmres_max_iters >= 1, "batched solve: gmres_max_iters must be >= 1");
    require(p.max_nl >= 0, "negative iteration limits");
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < B_; ++i) {
        require(x[i] != nullptr, "null iterate");
        out[i] = numa_krylov_result{};
    }
    if (!tuned_) autotune();
    const std::vector<char> all(B_, 1);
    set_active(all);
    for (auto& I : inst_) {
        I.xc = I.xa;
        I.xt = I.x2;
        I.Fc = I.Fx;
        I.Ft = I.Fn;
    }
    // the instances iterate in their own buffers (argument tables independent
    // of the caller's arrays); the caller's uploads precede this stream's work
    copy_x([&](Inst& I) { return (const double*)x[idx(I)]; }, [](Inst& I) { return I.xa; });
    eval_F(false);
    read_rows(MISC(inst_[0], S_FNORM), 1);
    std::vector<double> rnorm(B_), rprev(B_), eta_prev(B_), res0(B_), eta1(B_, 0.0), fn(B_, 0.0),
        lam(B_, 1.0);
    std::vector<char> running(B_, 1);
    for (size_t i = 0; i < B_; ++i) {
        out[i].n_F = 1;
        rnorm[i] = res0[i] = rprev[i] = hs_[i];
        out[i].res0 = res0[i];
        eta_prev[i] = p.eta_init;
    }
    for (int64_t k = 0; k < p.max_nl; ++k) {
        bool any_run = false;
        for (size_t i = 0; i < B_; ++i) {
            if (!running[i]) continue;
            if (nl_converged(p, rnorm[i], res0[i], out[i].iters, rprev[i])) {
                out[i].converged = 1;
                running[i] = 0;
                continue;
            }
            eta1[i] = forcing(p, k, rnorm[i], rprev[i], eta_prev[i], res0[i]);
            any_run = true;
        }
        if (!any_run) break;
        set_active(running);
        point_setup();
        gmres_batch(running, eta1, p.gmres_restart, p.gmres_max_iters);
        std::vector<char> ls(B_, 0), ok(B_, 0);
        bool any_ls = false;
        for (size_t i = 0; i < B_; ++i) {
            if (!running[i]) continue;
            const numa_gmres_result& g = gst_[i].res;
            out[i].n_JVP += g.matvecs;
            out[i].n_lin_iters += g.iters;
            out[i].n_prec += g.precs;
            if (!g.converged) ++out[i].n_lin_failures;
            if (g.breakdown) {
                running[i] = 0;
                continue;
            }
            lam[i] = 1.0;
            ls[i] = lam[i] >= p.min_lambda ? 1 : 0;
            any_ls = any_ls || ls[i];
        }
        // Armijo backtracking x_t = x_c - lam z (krylov_newton.hpp:96-108)
        while (any_ls) {
            for (size_t i = 0; i < B_; ++i) hup_[i] = ls[i] ? lam[i] : 0.0;
            write_rows(MISC(inst_[0], S_LSA), 1);
            set_active(ls);
            eval_F(true);
            read_rows(MISC(inst_[0], S_FNORM), 1);
            any_ls = false;
            for (size_t i = 0; i < B_; ++i) {
                if (!ls[i]) continue;
                ++out[i].n_F;
                fn[i] = hs_[i];
                if (fn[i] <= (1.0 - p.armijo * lam[i]) * rnorm[i]) {
                    ok[i] = 1;
                    ls[i] = 0;
                    continue;
                }
                lam[i] *= 0.5;
                if (!(lam[i] >= p.min_lambda)) ls[i] = 0;
                any_ls = any_ls || ls[i];
            }
        }
        for (size_t i = 0; i < B_; ++i) {
            if (!running[i]) continue;
            if (!ok[i]) {
                running[i] = 0;
                continue;
            }
            Inst& I = inst_[i];
            std::swap(I.xc, I.xt);
            std::swap(I.Fc, I.Ft);
            rprev[i] = rnorm[i];
            rnorm[i] = fn[i];
            eta_prev[i] = eta1[i];
            ++out[i].iters;
        }
    }
    for (size_t i = 0; i < B_; ++i) {
        if (rnorm[i] <= p.rtol_nl * res0[i]) out[i].converged = 1;
        out[i].res = rnorm[i];
    }
    set_active(all);
    copy_x([](Inst& I) { return (const double*)I.xc; }, [&](Inst& I) { return x[idx(I)]; });
    sync();
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (size_t i = 0; i < B_; ++i) out[i].t_total = t;
}

}  // namespace numa_hip

struct numa_hip_batch {
    numa_hip::Ctx* ctx;
    std::unique_ptr<numa_hip::BatchSolver> impl;
};

using numa_hip::require;

extern "C" {

int numa_hip_batch_create(numa_hip_ctx* ctx, int64_t count, const int64_t* n,
                          const double* lambda, const numa_solver_config* cfg,
                          int64_t gmres_restart, numa_hip_batch** out) {
    NUMA_API_BEGIN
    require(ctx && ctx->impl && cfg && out && n && lambda, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(ctx->impl->device));
    auto h = std::make_unique<numa_hip_batch>();
    h->ctx = ctx->impl;
    h->impl = std::make_unique<numa_hip::BatchSolver>(*ctx->impl, count, n, lambda, *cfg,
                                                      gmres_restart);
    ++ctx->impl->live_handles;
    *out = h.release();
    NUMA_API_END
}

int numa_hip_batch_destroy(numa_hip_batch* h) {
    NUMA_API_BEGIN
    if (h) {
        NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
        --h->ctx->live_handles;
        delete h;
    }
    NUMA_API_END
}

int numa_hip_batch_set_graphs(numa_hip_batch* h, int64_t on) {
    NUMA_API_BEGIN
    require(h != nullptr, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    h->impl->set_graphs(on != 0);
    NUMA_API_END
}

int numa_hip_batch_set_smoothing(numa_hip_batch* h, int64_t count, const int64_t* form) {
    NUMA_API_BEGIN
    require(h != nullptr, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    h->impl->set_smoothing(count, form);
    NUMA_API_END
}

int numa_hip_batch_solve(numa_hip_batch* h, const numa_krylov_params* p, double* const* x_dev,
                         numa_krylov_result* results, double* seconds) {
    NUMA_API_BEGIN
    require(h && p && x_dev && results, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    // the caller's uploads on the context stream precede the batch's reads
    const auto t0 = std::chrono::steady_clock::now();
    h->impl->solve(*p, x_dev, results);
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    NUMA_API_END
}

int numa_hip_batch_stats(numa_hip_batch* h, int64_t reset, numa_batch_stats* out) {
    NUMA_API_BEGIN
    require(h && out, "null argument");
    *out = h->impl->stats();
    if (reset) h->impl->reset_stats();
    NUMA_API_END
}

}  // extern "C"
