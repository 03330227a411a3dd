// sweep.hip - batched sweep over many Bratu3D instances of different grid
// size and lambda on one GPU.
//
// A sweep handle owns W workers; each worker is a HIP stream (its own
// context: stream, reduction scratch, scalar slots) with one device-resident
// solver per grid size, created at first use and reused for every instance of
// that size (set_lambda).  A run hands the instances, largest grids first, to
// the workers through an atomic counter; W host threads drive the workers so
// that up to W instances advance concurrently and their kernels overlap on
// the device.  With graphs on, every Arnoldi step of a worker's solver is a
// hipGraph launch, captured once per step index and reused across instances,
// so a Krylov step costs one launch instead of ~10^2.
//
// The instances of a run share the solver configuration and Newton
// parameters; their iterates are device vectors owned by the caller and are
// overwritten with the solutions.  Each instance's result is the one of the
// same solver run alone: the kernels are deterministic and per stream.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <thread>

#include "solver_base.hpp"

namespace numa_hip {

struct SweepWorker {
    std::unique_ptr<Ctx> ctx;
    std::map<int64_t, std::unique_ptr<SolverBase>> solvers;
};

class Sweep {
public:
    Sweep(Ctx& main, int64_t workers, const numa_solver_config& cfg, int64_t restart)
        : main_(main), cfg_(cfg), restart_(restart) {
        require(workers >= 1 && workers <= 32, "workers must be in [1, 32]");
        require(restart >= 1 && restart <= 256, "gmres restart must be in [1, 256]");
        require(cfg.precond != NUMA_PRECOND_LEARNED,
                "the sweep runs the V-cycle / Jacobi / no preconditioner");
        for (int64_t w = 0; w < workers; ++w) {
            SweepWorker wk;
            wk.ctx = std::make_unique<Ctx>(main.device);
            workers_.push_back(std::move(wk));
        }
    }

    void run(int64_t count, const int64_t* n, const double* lambda, double* const* x,
             const numa_krylov_params& p, bool graphs, numa_krylov_result* results,
             double* seconds) {
        require(count >= 0, "negative instance count");
        for (int64_t i = 0; i < count; ++i) {
            require(n[i] >= 2 && n[i] <= 1024, "instance grid size must be in [2, 1024]");
            require(std::isfinite(lambda[i]), "lambda must be finite");
            require(x[i] != nullptr, "null iterate");
        }
        // caller's uploads on the main stream precede the workers' reads
        main_.sync();
        std::vector<int64_t> order((size_t)count);
        for (int64_t i = 0; i < count; ++i) order[(size_t)i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t a, int64_t b) { return n[a] > n[b]; });
        std::atomic<int64_t> next{0};
        std::atomic<bool> failed{false};
        std::exception_ptr err;
        std::mutex err_mu;
        const auto t0 = std::chrono::steady_clock::now();
        auto work = [&](SweepWorker& wk) {
            try {
                NUMA_HIP_CHECK(hipSetDevice(wk.ctx->device));
                while (!failed.load()) {
                    const int64_t q = next.fetch_add(1);
                    if (q >= count) break;
                    const int64_t i = order[(size_t)q];
                    auto& sv = wk.solvers[n[i]];
                    if (!sv) sv.reset(make_bratu_solver(*wk.ctx, n[i], lambda[i], cfg_, restart_));
                    sv->set_lambda(lambda[i]);
                    sv->set_graphs(graphs);
                    sv->solve(p, x[i], results[i]);
                }
            } catch (...) {
                std::lock_guard<std::mutex> g(err_mu);
                if (!err) err = std::current_exception();
                failed = true;
            }
        };
        std::vector<std::thread> threads;
        for (size_t w = 1; w < workers_.size(); ++w)
            threads.emplace_back(work, std::ref(workers_[w]));
        work(workers_[0]);
        for (auto& t : threads) t.join();
        for (auto& wk : workers_) wk.ctx->sync();
        if (seconds)
            *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (err) std::rethrow_exception(err);
    }

    int64_t workers() const { return (int64_t)workers_.size(); }
    Ctx& main() { return main_; }

private:
    Ctx& main_;
    numa_solver_config cfg_;
    int64_t restart_;
    std::vector<SweepWorker> workers_;
};

}  // namespace numa_hip

struct numa_hip_sweep {
    numa_hip::Ctx* ctx;
    std::unique_ptr<numa_hip::Sweep> impl;
};

using numa_hip::require;

extern "C" {

int numa_hip_sweep_create(numa_hip_ctx* ctx, int64_t workers, const numa_solver_config* cfg,
                          int64_t gmres_restart, numa_hip_sweep** out) {
    NUMA_API_BEGIN
    require(ctx && ctx->impl && cfg && out, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(ctx->impl->device));
    auto h = std::make_unique<numa_hip_sweep>();
    h->ctx = ctx->impl;
    h->impl = std::make_unique<numa_hip::Sweep>(*ctx->impl, workers, *cfg, gmres_restart);
    ++ctx->impl->live_handles;
    *out = h.release();
    NUMA_API_END
}

int numa_hip_sweep_destroy(numa_hip_sweep* h) {
    NUMA_API_BEGIN
    if (h) {
        NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
        --h->ctx->live_handles;
        delete h;
    }
    NUMA_API_END
}

int numa_hip_sweep_run(numa_hip_sweep* h, int64_t count, const int64_t* n, const double* lambda,
                       double* const* x_dev, const numa_krylov_params* p, int64_t graphs,
                       numa_krylov_result* results, double* seconds) {
    NUMA_API_BEGIN
    require(h && p && results && (count == 0 || (n && lambda && x_dev)), "null argument");
    require(p->max_nl >= 0 && p->gmres_max_iters >= 0 && p->gmres_restart >= 0,
            "negative iteration limits");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    h->impl->run(count, n, lambda, x_dev, *p, graphs != 0, results, seconds);
    NUMA_API_END
}

// Instances of a sweep from a seed: n uniformly from `sizes`, lambda ~
// U[lambda_lo, lambda_hi], with std::mt19937_64 and U[0,1) = (u64 >> 11) 2^-53
// (size index = u64 % nsizes, drawn before lambda, per instance).
int numa_hip_sweep_instances(int64_t seed, int64_t count, const int64_t* sizes, int64_t nsizes,
                             double lambda_lo, double lambda_hi, int64_t* n_out,
                             double* lambda_out) {
    NUMA_API_BEGIN
    require(count >= 0 && nsizes >= 1 && sizes && n_out && lambda_out, "bad argument");
    require(lambda_lo <= lambda_hi, "lambda_lo > lambda_hi");
    std::mt19937_64 rng((uint64_t)seed);
    for (int64_t i = 0; i < count; ++i) {
        n_out[i] = sizes[rng() % (uint64_t)nsizes];
        lambda_out[i] = lambda_lo + (lambda_hi - lambda_lo) * ((double)(rng() >> 11) * 0x1.0p-53);
    }
    NUMA_API_END
}

// out[i] = scale * U[0,1) from std::mt19937_64(seed), U = (u64 >> 11) 2^-53.
// Host memory, no device work (the seeded starting points of a sweep).
int numa_hip_uniform_host(int64_t seed, int64_t count, double scale, double* out) {
    NUMA_API_BEGIN
    require(count >= 0 && (count == 0 || out), "bad argument");
    std::mt19937_64 rng((uint64_t)seed);
    for (int64_t i = 0; i < count; ++i) out[i] = scale * ((double)(rng() >> 11) * 0x1.0p-53);
    NUMA_API_END
}

}  // extern "C"
