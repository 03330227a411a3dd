// solver_base.hpp - the type-erased interface of the device-resident Bratu3D
// solver, shared by the solver and the learned-preconditioner translation
// units.
#pragma once

#include <functional>
#include <memory>

#include "common.hpp"
#include "mg.hpp"

namespace numa_hip {

// A learned preconditioner z = M(r) on an n^3 grid (FP32 or FP64 vectors on
// the context's stream).
struct NetApply {
    virtual ~NetApply() = default;
    virtual void apply(int64_t n, const float* r, float* z) = 0;
    virtual void apply(int64_t n, const double* r, double* z) = 0;
};

// Called after every preconditioner application with the device input and
// output vectors (T = the solver's Krylov type), in stream order.
using CaptureFn = std::function<void(const void* in, const void* out)>;

struct LearnedStats {
    int64_t net_applications = 0;  // preconditioner calls served by the network
    int64_t gmg_applications = 0;  // calls served by the V-cycle
    int64_t fallbacks = 0;         // linear solves switched to the V-cycle
};

struct SolverBase {
    virtual ~SolverBase() = default;
    virtual numa_gmres_result gmres_entry(const double* x0, const double* b, double* z, double rtol,
                                          int64_t restart, int64_t max_iters, bool zero_guess,
                                          const double* r_init = nullptr, bool save = false) = 0;
    // The Arnoldi relation saved by the last gmres_entry(save): returns k and
    // writes Hbar ((k+1) x k, row major) and V_0..V_k (FP64 device vectors)
    // where the pointers are not null.
    virtual int64_t saved_space(double* hbar, double* const* V) = 0;
    // gmres_recycled (gmres.hpp:229-323) at x0 on the given relation (k, Hbar,
    // V_0..V_k device vectors), or on the saved one when V is null; save keeps
    // the last cycle of its GMRES.
    virtual numa_gmres_result gmres_recycled_entry(const double* x0, const double* b, double* z,
                                                   double rtol, int64_t restart,
                                                   int64_t max_iters, int64_t k,
                                                   const double* hbar, const double* const* V,
                                                   bool save) = 0;
    virtual void solve(const numa_krylov_params& p, double* x, numa_krylov_result& out) = 0;
    virtual int64_t size() const = 0;
    virtual int64_t grid_n() const = 0;
    virtual double primitive_cost(int64_t prim, int64_t reps, double* bytes) = 0;
    virtual void set_lambda(double lambda) = 0;
    virtual void set_capture(CaptureFn fn) = 0;
    // Attaches a learned preconditioner (NUMA_PRECOND_LEARNED solvers): the
    // network serves each linear solve until a step reduces the FGMRES
    // residual estimate by less than the factor `ratio`; the rest of that
    // solve uses the V-cycle.
    virtual void set_learned(std::shared_ptr<NetApply> net, double ratio) = 0;
    virtual LearnedStats learned_stats() const = 0;
    // The multigrid preconditioner (null without one).
    virtual const MGBase* multigrid() const = 0;
    virtual MGBase* multigrid() = 0;
    // Replay each Arnoldi step of solve() from a captured hipGraph (one graph
    // per step index and iterate buffers, captured at first use).  Solvers with
    // the learned preconditioner or a capture hook run eagerly.
    virtual void set_graphs(bool on) = 0;
    virtual int64_t graphs_cached() const = 0;
    // Device time per primitive (Timed / TimedPrecond, operators/timing.hpp);
    // timing runs the Arnoldi steps eagerly.
    virtual void set_timing(bool on) = 0;
    virtual numa_timers timers(bool reset) = 0;
    // tensor_free_curvature_2018 at (x, d) into out (device FP64 vectors,
    // out distinct from x and d); returns the step h.
    virtual double tf2018_curvature(const double* x, const double* d, double* out) = 0;
    // inverse_smallest_eigenpair of J(x) (analysis/spectral.hpp:34-115);
    // vec (device, may be null) receives the last iterate.
    virtual numa_eigen_result eigen_smallest(const double* x, const numa_eigen_params& p,
                                             double* vec) = 0;
    // NaN into the `call`-th F evaluation (what 0) or J v product (what 1)
    // from now on; fires once; call = 0 disarms.
    virtual void inject_nan(int64_t what, int64_t call) = 0;
};

SolverBase* make_bratu_solver(Ctx& c, int64_t n, double lambda, const numa_solver_config& cfg,
                              int64_t restart);

}  // namespace numa_hip

// Opaque C handle of a solver.
struct numa_hip_solver {
    numa_hip::Ctx* ctx;
    std::unique_ptr<numa_hip::SolverBase> impl;
};
