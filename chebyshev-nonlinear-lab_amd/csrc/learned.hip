// learned.hip - the learned preconditioner: network state and optimiser,
// training on (residual, V-cycle correction) pairs that the solver generates
// while it runs, inference as an FGMRES preconditioner, and the C entry points.
//
// Preconditioner form: z = M(r) = rho / (n+1)^2 * net(r / rho) with
// rho = rms(r), so M is positively homogeneous (M(c r) = c M(r), c > 0) like
// the V-cycle it imitates, and one network serves every grid size: the
// network works in grid units, the factor (n+1)^-2 restores the scale of the
// operator (n+1)^2 (6u - sum of neighbours).
//
// Training pair from a preconditioner call r -> z_V (the V-cycle's output):
//   input x = r / rho, target y = (n+1)^2 z_V / rho, and the loss of a sample
//   NUMA_NET_LOSS_CORRECTION: ||net(x) - y||^2 / ||y||^2
//   NUMA_NET_LOSS_RESIDUAL:   ||A7 (net(x) - y)||^2 / ||A7 y||^2,
//   A7 u = 6u - (sum of neighbours), averaged over the batch.  The residual
//   form measures the error where a preconditioner is judged (the operator's
//   residual, as a sparse approximate inverse is fitted): the smooth part of
//   the V-cycle's correction, which scales with (n+1)^2 and which no
//   translation-invariant network of 13^3 reach can reproduce for every n,
//   does not dominate it.
#include <chrono>
#include <cmath>
#include <map>
#include <random>

#include "convnet.hpp"
#include "solver_base.hpp"

namespace numa_hip {

using cnn::bf16x8;

namespace {

constexpr int kMaxBatch = 64;
constexpr int kWgradWorkers = 256;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void ensure(size_t b) {
        if (b <= bytes) return;
        if (p) NUMA_HIP_CHECK(hipFree(p));
        p = nullptr;
        bytes = 0;
        NUMA_HIP_CHECK(hipMalloc(&p, b));
        bytes = b;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

double uniform01(std::mt19937_64& rng) { return (double)(rng() >> 11) * 0x1.0p-53; }

}  // namespace

class ConvNet : public NetApply {
public:
    ConvNet(Ctx& c, uint64_t seed) : c_(c) {
        params_.ensure(cnn::kParams * sizeof(float));
        grads_.ensure(cnn::kParams * sizeof(float));
        adam_m_.ensure(cnn::kParams * sizeof(float));
        adam_v_.ensure(cnn::kParams * sizeof(float));
        for (int l = 0; l < 4; ++l) {
            wpk_[l].ensure(cnn::kSteps * 64 * sizeof(bf16x8));
            wpkT_[l].ensure(cnn::kSteps * 64 * sizeof(bf16x8));
        }
        small_.ensure(4 * kMaxBatch * sizeof(float));
        loss_.ensure((kMaxBatch + 1) * sizeof(double));
        part_.ensure((size_t)kWgradWorkers * 27 * cnn::kC * cnn::kC * sizeof(float));
        bpart_.ensure((size_t)kWgradWorkers * cnn::kC * sizeof(float));
        std::vector<float> ones(kMaxBatch, 1.0f);
        NUMA_HIP_CHECK(hipMemcpy(ones_(), ones.data(), kMaxBatch * sizeof(float),
                                 hipMemcpyHostToDevice));
        set_params(init_params(seed).data());
    }

    // He-normal initialisation (std sqrt(2 / fan_in); sqrt(1 / fan_in) for the
    // linear output layer), zero biases, from mt19937_64 + Box-Muller.
    static std::vector<float> init_params(uint64_t seed) {
        std::vector<float> w((size_t)cnn::kParams, 0.0f);
        std::mt19937_64 rng(seed);
        auto normal = [&]() {
            const double u1 = 1.0 - uniform01(rng), u2 = uniform01(rng);
            return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
        };
        auto fill = [&](int64_t off, int64_t count, double stdv) {
            for (int64_t i = 0; i < count; ++i) w[(size_t)(off + i)] = (float)(stdv * normal());
        };
        fill(cnn::kW0, cnn::kC * cnn::kTaps, std::sqrt(2.0 / 27.0));
        for (int l = 0; l < 4; ++l) fill(cnn::mid_w(l), cnn::kMidW, std::sqrt(2.0 / (24.0 * 27.0)));
        fill(cnn::kW5, cnn::kC * cnn::kTaps, std::sqrt(1.0 / (24.0 * 27.0)));
        return w;
    }

    void set_params(const float* host) {
        NUMA_HIP_CHECK(hipMemcpyAsync(params_.p, host, cnn::kParams * sizeof(float),
                                      hipMemcpyHostToDevice, c_.stream));
        NUMA_HIP_CHECK(hipMemsetAsync(adam_m_.p, 0, cnn::kParams * sizeof(float), c_.stream));
        NUMA_HIP_CHECK(hipMemsetAsync(adam_v_.p, 0, cnn::kParams * sizeof(float), c_.stream));
        adam_t_ = 0;
        pack();
        c_.sync();
    }
    void get(const DevBuf& b, float* host) {
        NUMA_HIP_CHECK(hipMemcpyAsync(host, b.p, cnn::kParams * sizeof(float),
                                      hipMemcpyDeviceToHost, c_.stream));
        c_.sync();
    }
    void get_params(float* host) { get(params_, host); }
    void get_grads(float* host) { get(grads_, host); }

    // Inference: z = rho / (n+1)^2 * net(r / rho) (batch 1).
    void apply(int64_t n, const float* r, float* z) override { apply_t(n, r, z); }
    void apply(int64_t n, const double* r, double* z) override { apply_t(n, r, z); }

    // Raw network on a batch (no scaling): out = net(x).
    void forward(int64_t n, int B, const float* x, float* out) {
        check_shape(n, B);
        infer(n, B, x, ones_(), ones_(), out);
    }

    // Loss and gradients of a batch; normalized = training-pair scaling of
    // (r, z_V) as described at the top of the file, else (x, y) are used as
    // given.  The mean loss goes to *mean_dev (device).
    void loss_grad(int64_t n, int B, const float* x, const float* z, bool normalized,
                   int64_t loss_kind, double* mean_dev);

    void adam(double lr) {
        ++adam_t_;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = (float)(1.0 - std::pow(0.9, (double)adam_t_));
        const float bc2 = (float)(1.0 - std::pow(0.999, (double)adam_t_));
        hipLaunchKernelGGL(cnn::k_adam, dim3((unsigned)((cnn::kParams + 255) / 256)), dim3(256), 0,
                           c_.stream, cnn::kParams, params_.as<float>(), adam_m_.as<float>(),
                           adam_v_.as<float>(), grads_.as<float>(), (float)lr, b1, b2, 1e-8f, bc1,
                           bc2);
        NUMA_HIP_LAUNCH_CHECK();
        pack();
    }

    double read_mean() {
        double v = 0.0;
        NUMA_HIP_CHECK(hipMemcpyAsync(&v, loss_.as<double>() + kMaxBatch, sizeof(double),
                                      hipMemcpyDeviceToHost, c_.stream));
        c_.sync();
        return v;
    }
    double* mean_slot() { return loss_.as<double>() + kMaxBatch; }

    void train_from_solver(const numa_net_train_config& cfg, double* hist, int64_t hist_len,
                           numa_net_train_stats& st);

    // Average device ms of one inference (what = 0) or one training step of
    // `batch` samples (what = 1) at n^3, on synthetic input.
    double cost(int64_t n, int64_t what, int64_t batch, int64_t reps);

    Ctx& ctx() { return c_; }

private:
    float* ones_() { return small_.as<float>(); }
    float* xs_() { return small_.as<float>() + kMaxBatch; }
    float* ys_() { return small_.as<float>() + 2 * kMaxBatch; }
    float* coef_() { return small_.as<float>() + 3 * kMaxBatch; }
    float* P(int64_t off) { return params_.as<float>() + off; }
    float* G(int64_t off) { return grads_.as<float>() + off; }

    static void check_shape(int64_t n, int B) {
        require(n >= 1, "grid size must be positive");
        require(n <= 1024, "grid too large for the network workspace", NUMA_ERR_OVERFLOW);
        require(B >= 1 && B <= kMaxBatch, "batch must be in [1, 64]");
    }

    void pack() {
        for (int l = 0; l < 4; ++l) {
            hipLaunchKernelGGL(cnn::k_pack_mid<false>, dim3(cnn::kSteps), dim3(64), 0, c_.stream,
                               P(cnn::mid_w(l)), wpk_[l].as<bf16x8>());
            hipLaunchKernelGGL(cnn::k_pack_mid<true>, dim3(cnn::kSteps), dim3(64), 0, c_.stream,
                               P(cnn::mid_w(l)), wpkT_[l].as<bf16x8>());
        }
        NUMA_HIP_LAUNCH_CHECK();
    }

    static int sum_blocks(int64_t N3) { return (int)std::min<int64_t>(256, (N3 + 1023) / 1024); }

    template <class T>
    void apply_t(int64_t n, const T* r, T* z) {
        check_shape(n, 1);
        const int64_t N3 = n * n * n;
        const int nblk = sum_blocks(N3);
        hipLaunchKernelGGL((cnn::k_sample_sums<T, 0>), dim3(nblk, 1), dim3(256), 0, c_.stream, N3, r,
                           (const float*)nullptr, (const float*)nullptr, c_.red.partials);
        hipLaunchKernelGGL(cnn::k_sample_finish<0>, dim3(1), dim3(64), 0, c_.stream, nblk, 1, N3,
                           (n + 1.0) * (n + 1.0), (const double*)c_.red.partials, xs_(), ys_(),
                           (double*)nullptr, (double*)nullptr);
        NUMA_HIP_LAUNCH_CHECK();
        infer(n, 1, r, xs_(), ys_(), z);
    }

    template <class T>
    void infer(int64_t n, int B, const T* x, const float* xs, const float* os, T* out) {
        const int64_t N3 = n * n * n;
        const size_t abytes = (size_t)B * N3 * 3 * sizeof(bf16x8);
        ia_.ensure(abytes);
        ib_.ensure(abytes);
        bf16x8* cur = ia_.as<bf16x8>();
        bf16x8* nxt = ib_.as<bf16x8>();
        const int ni = (int)n;
        hipLaunchKernelGGL((cnn::k_conv_first<T, false>), dim3((unsigned)((N3 + 255) / 256), B),
                           dim3(256), 0, c_.stream, ni, x, xs, (const float*)P(cnn::kW0),
                           (const float*)P(cnn::kB0), cur, (bf16x8*)nullptr);
        for (int l = 0; l < 4; ++l) {
            hipLaunchKernelGGL(cnn::k_conv24<cnn::CE_INFER>, dim3(cnn::kConvWorkers),
                               dim3(cnn::kConvThreads), 0, c_.stream, ni, B, (const bf16x8*)cur,
                               (const bf16x8*)wpk_[l].as<bf16x8>(), (const float*)P(cnn::mid_b(l)),
                               (const bf16x8*)nullptr, nxt, (bf16x8*)nullptr);
            std::swap(cur, nxt);
        }
        hipLaunchKernelGGL(cnn::k_conv_last<T>, dim3(cnn::kWorkers), dim3(256), 0, c_.stream, ni, B,
                           (const bf16x8*)cur, (const float*)P(cnn::kW5), (const float*)P(cnn::kB5),
                           os, out);
        NUMA_HIP_LAUNCH_CHECK();
    }

    Ctx& c_;
    DevBuf params_, grads_, adam_m_, adam_v_, small_, loss_, part_, bpart_;
    DevBuf wpk_[4], wpkT_[4];
    DevBuf ia_, ib_;                 // inference ping-pong activations
    DevBuf act_[5], pre_[5], ga_, gb_, out_, g6_, e_;  // training workspace
    int64_t adam_t_ = 0;
};

void ConvNet::loss_grad(int64_t n, int B, const float* x, const float* z, bool normalized,
                        int64_t loss_kind, double* mean_dev) {
    check_shape(n, B);
    require(loss_kind == NUMA_NET_LOSS_CORRECTION || loss_kind == NUMA_NET_LOSS_RESIDUAL,
            "unknown loss");
    const int64_t N3 = n * n * n;
    const int ni = (int)n;
    const size_t abytes = (size_t)B * N3 * 3 * sizeof(bf16x8);
    for (int l = 0; l < 5; ++l) {
        act_[l].ensure(abytes);
        pre_[l].ensure(abytes);
    }
    ga_.ensure(abytes);
    gb_.ensure(abytes);
    out_.ensure((size_t)B * N3 * sizeof(float));
    g6_.ensure((size_t)B * N3 * sizeof(float));
    if (loss_kind == NUMA_NET_LOSS_RESIDUAL) e_.ensure((size_t)B * N3 * sizeof(float));
    const int nblk = sum_blocks(N3);
    require(2 * B * nblk <= kMaxPartials, "batch too large for the reduction scratch");
    const hipStream_t s = c_.stream;
    // weight-gradient workgroups: one per tile up to kWgradWorkers (multiple of 8)
    const int64_t wtiles = cnn::Tiles<>((int)n, B).total;
    const int WG = (int)std::min<int64_t>(kWgradWorkers, (wtiles + 7) / 8 * 8);
    const float* xs = ones_();
    const float* ys = ones_();
    if (normalized) {
        hipLaunchKernelGGL((cnn::k_sample_sums<float, 0>), dim3(nblk, B), dim3(256), 0, s, N3, x,
                           (const float*)nullptr, (const float*)nullptr, c_.red.partials);
        hipLaunchKernelGGL(cnn::k_sample_finish<2>, dim3(1), dim3(64), 0, s, nblk, B, N3,
                           (n + 1.0) * (n + 1.0), (const double*)c_.red.partials, xs_(), ys_(),
                           (double*)nullptr, (double*)nullptr);
        xs = xs_();
        ys = ys_();
    }
    // forward, keeping pre-activations and activations
    hipLaunchKernelGGL((cnn::k_conv_first<float, true>), dim3((unsigned)((N3 + 255) / 256), B),
                       dim3(256), 0, s, ni, x, xs, (const float*)P(cnn::kW0),
                       (const float*)P(cnn::kB0), act_[0].as<bf16x8>(), pre_[0].as<bf16x8>());
    for (int l = 0; l < 4; ++l)
        hipLaunchKernelGGL(cnn::k_conv24<cnn::CE_TRAIN>, dim3(cnn::kConvWorkers),
                           dim3(cnn::kConvThreads), 0, s, ni, B, (const bf16x8*)act_[l].as<bf16x8>(),
                           (const bf16x8*)wpk_[l].as<bf16x8>(), (const float*)P(cnn::mid_b(l)),
                           (const bf16x8*)nullptr, act_[l + 1].as<bf16x8>(), pre_[l + 1].as<bf16x8>());
    hipLaunchKernelGGL(cnn::k_conv_last<float>, dim3(cnn::kWorkers), dim3(256), 0, s, ni, B,
                       (const bf16x8*)act_[4].as<bf16x8>(), (const float*)P(cnn::kW5),
                       (const float*)P(cnn::kB5), (const float*)ones_(), out_.as<float>());
    NUMA_HIP_LAUNCH_CHECK();
    // loss, dL/dout
    if (loss_kind == NUMA_NET_LOSS_CORRECTION)
        hipLaunchKernelGGL((cnn::k_sample_sums<float, 1>), dim3(nblk, B), dim3(256), 0, s, N3, z,
                           (const float*)out_.as<float>(), ys, c_.red.partials);
    else
        hipLaunchKernelGGL(cnn::k_res_sums<float>, dim3(nblk, B), dim3(256), 0, s, ni,
                           (const float*)out_.as<float>(), z, ys, e_.as<float>(), c_.red.partials);
    hipLaunchKernelGGL(cnn::k_sample_finish<1>, dim3(1), dim3(64), 0, s, nblk, B, N3, 0.0,
                       (const double*)c_.red.partials, coef_(), (float*)nullptr,
                       loss_.as<double>(), mean_dev);
    if (loss_kind == NUMA_NET_LOSS_CORRECTION)
        hipLaunchKernelGGL(cnn::k_out_grad<float>,
                           dim3((unsigned)std::min<int64_t>(1024, (N3 + 255) / 256), B), dim3(256),
                           0, s, N3, (const float*)out_.as<float>(), z, ys, (const float*)coef_(),
                           g6_.as<float>());
    else
        hipLaunchKernelGGL(cnn::k_res_grad, dim3((unsigned)((N3 + 255) / 256), B), dim3(256), 0, s,
                           ni, (const float*)e_.as<float>(), (const float*)coef_(), g6_.as<float>());
    NUMA_HIP_LAUNCH_CHECK();
    // last layer
    hipLaunchKernelGGL(cnn::k_wgrad_last, dim3(WG), dim3(256), 0, s, ni, B,
                       (const bf16x8*)act_[4].as<bf16x8>(), (const float*)g6_.as<float>(),
                       part_.as<float>());
    hipLaunchKernelGGL(cnn::k_reduce_partials<1>, dim3(cnn::reduce_blocks<1>()), dim3(256), 0, s,
                       WG, (const float*)part_.as<float>(), G(cnn::kW5));
    hipLaunchKernelGGL(cnn::k_bwd_last, dim3((unsigned)((N3 + 255) / 256), B), dim3(256), 0, s, ni,
                       (const float*)g6_.as<float>(), (const float*)P(cnn::kW5),
                       (const bf16x8*)pre_[4].as<bf16x8>(), ga_.as<bf16x8>());
    NUMA_HIP_LAUNCH_CHECK();
    bf16x8* g = ga_.as<bf16x8>();
    bf16x8* gn = gb_.as<bf16x8>();
    for (int l = 3; l >= 0; --l) {
        hipLaunchKernelGGL(cnn::k_wgrad_mid, dim3(WG, 3), dim3(256), 0, s, ni, B,
                           (const bf16x8*)g, (const bf16x8*)act_[l].as<bf16x8>(), part_.as<float>(),
                           bpart_.as<float>());
        hipLaunchKernelGGL(cnn::k_reduce_partials<2>, dim3(cnn::reduce_blocks<2>()), dim3(256), 0,
                           s, WG, (const float*)part_.as<float>(), G(cnn::mid_w(l)));
        hipLaunchKernelGGL(cnn::k_reduce_partials<3>, dim3(cnn::reduce_blocks<3>()), dim3(256), 0,
                           s, WG, (const float*)bpart_.as<float>(), G(cnn::mid_b(l)));
        hipLaunchKernelGGL(cnn::k_conv24<cnn::CE_BWD>, dim3(cnn::kConvWorkers),
                           dim3(cnn::kConvThreads), 0, s, ni, B, (const bf16x8*)g,
                           (const bf16x8*)wpkT_[l].as<bf16x8>(), (const float*)nullptr,
                           (const bf16x8*)pre_[l].as<bf16x8>(), gn, (bf16x8*)nullptr);
        NUMA_HIP_LAUNCH_CHECK();
        std::swap(g, gn);
    }
    hipLaunchKernelGGL(cnn::k_wgrad_first<float>, dim3(WG), dim3(256), 0, s, ni, B, x, xs,
                       (const bf16x8*)g, part_.as<float>());
    hipLaunchKernelGGL(cnn::k_reduce_partials<0>, dim3(cnn::reduce_blocks<0>()), dim3(256), 0, s,
                       WG, (const float*)part_.as<float>(), G(cnn::kW0));
    NUMA_HIP_LAUNCH_CHECK();
}

// In-solver data generator: mixed-precision Newton-FGMRES solves with the
// V-cycle preconditioner on random instances; every preconditioner call
// (r, z_V) becomes a training pair, copied into the batch buffer on the
// stream and consumed by an optimiser step when the batch is full.  Pairs
// are never stored beyond one batch.
void ConvNet::train_from_solver(const numa_net_train_config& cfg, double* hist, int64_t hist_len,
                                numa_net_train_stats& st) {
    require(cfg.pairs > 0, "pairs must be positive");
    require(cfg.batch >= 1 && cfg.batch <= kMaxBatch, "batch must be in [1, 64]");
    require(cfg.lr > 0.0, "learning rate must be positive");
    require(cfg.lambda_lo <= cfg.lambda_hi, "lambda_lo > lambda_hi");
    require(cfg.gmres_restart >= 1 && cfg.gmres_restart <= 256, "gmres restart must be in [1, 256]");
    require(cfg.loss == NUMA_NET_LOSS_CORRECTION || cfg.loss == NUMA_NET_LOSS_RESIDUAL,
            "unknown loss");
    std::vector<int64_t> sizes;
    for (int64_t s : cfg.sizes)
        if (s > 0) sizes.push_back(s);
    require(!sizes.empty(), "no grid sizes given");
    for (int64_t s : sizes) require(s >= 2 && s <= 512, "training grid sizes must be in [2, 512]");

    const auto t0 = std::chrono::steady_clock::now();
    st = numa_net_train_stats{};
    std::mt19937_64 rng((uint64_t)cfg.seed);
    std::map<int64_t, std::unique_ptr<SolverBase>> solvers;
    numa_solver_config sc{};
    sc.method = NUMA_METHOD_NEWTON;
    sc.precond = NUMA_PRECOND_GMG;
    sc.precision = NUMA_PRECISION_MIXED;
    sc.flexible = 1;
    sc.jv_mode = NUMA_DERIV_FD;
    sc.d2_mode = NUMA_DERIV_FD;
    sc.mg = cfg.mg;
    const int64_t B = cfg.batch;
    DevBuf rb, zb, xd, histd;
    const int64_t hslots = std::max<int64_t>(hist_len, 1);
    histd.ensure((size_t)hslots * sizeof(double));
    int64_t fill = 0, ncur = 0;
    std::vector<double> x0;

    auto flush = [&]() {
        if (fill == 0) return;
        const int64_t N3 = ncur * ncur * ncur;
        double* slot = histd.as<double>() + std::min<int64_t>(st.steps, hslots - 1);
        loss_grad(ncur, (int)fill, rb.as<float>(), zb.as<float>(), true, cfg.loss, slot);
        // step size decays linearly to lr_end (lr/10 if 0) over the pairs
        const double lr_end = cfg.lr_end > 0.0 ? cfg.lr_end : 0.1 * cfg.lr;
        adam(cfg.lr + (lr_end - cfg.lr) * (double)(st.pairs - fill) / (double)cfg.pairs);
        (void)N3;
        ++st.steps;
        fill = 0;
    };

    while (st.pairs < cfg.pairs) {
        const int64_t n = sizes[(size_t)(rng() % sizes.size())];
        const double lambda = cfg.lambda_lo + (cfg.lambda_hi - cfg.lambda_lo) * uniform01(rng);
        const int64_t N3 = n * n * n;
        auto& sv = solvers[n];
        if (!sv) sv.reset(make_bratu_solver(c_, n, lambda, sc, cfg.gmres_restart));
        sv->set_lambda(lambda);
        x0.resize((size_t)N3);
        for (auto& v : x0) v = cfg.noise * uniform01(rng);
        xd.ensure((size_t)N3 * sizeof(double));
        rb.ensure((size_t)B * N3 * sizeof(float));
        zb.ensure((size_t)B * N3 * sizeof(float));
        NUMA_HIP_CHECK(hipMemcpyAsync(xd.p, x0.data(), (size_t)N3 * sizeof(double),
                                      hipMemcpyHostToDevice, c_.stream));
        ncur = n;
        sv->set_capture([&](const void* in, const void* out) {
            if (st.pairs >= cfg.pairs) return;
            NUMA_HIP_CHECK(hipMemcpyAsync(rb.as<float>() + fill * N3, in, (size_t)N3 * sizeof(float),
                                          hipMemcpyDeviceToDevice, c_.stream));
            NUMA_HIP_CHECK(hipMemcpyAsync(zb.as<float>() + fill * N3, out, (size_t)N3 * sizeof(float),
                                          hipMemcpyDeviceToDevice, c_.stream));
            ++fill;
            ++st.pairs;
            if (fill == B) flush();
        });
        numa_krylov_result res{};
        try {
            sv->solve(cfg.solve, xd.as<double>(), res);
        } catch (...) {
            sv->set_capture(nullptr);
            throw;
        }
        sv->set_capture(nullptr);
        flush();
        ++st.solves;
    }
    c_.sync();
    const int64_t nh = std::min<int64_t>(st.steps, hslots);
    std::vector<double> h((size_t)nh);
    if (nh > 0)
        NUMA_HIP_CHECK(hipMemcpy(h.data(), histd.p, (size_t)nh * sizeof(double), hipMemcpyDeviceToHost));
    if (hist)
        for (int64_t i = 0; i < std::min(nh, hist_len); ++i) hist[i] = h[(size_t)i];
    st.loss_first = nh > 0 ? h.front() : 0.0;
    st.loss_last = nh > 0 ? h.back() : 0.0;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double ConvNet::cost(int64_t n, int64_t what, int64_t batch, int64_t reps) {
    require(reps > 0, "reps must be positive");
    require(what == 0 || what == 1, "what must be 0 (inference) or 1 (training step)");
    const int B = what == 0 ? 1 : (int)batch;
    check_shape(n, B);
    const int64_t N3 = n * n * n;
    DevBuf r, z;
    r.ensure((size_t)B * N3 * sizeof(float));
    z.ensure((size_t)B * N3 * sizeof(float));
    // smooth synthetic input: r = sin(pi i h) sin(pi j h) sin(pi k h) via the
    // first layer's own bounds; any finite data will do for timing
    std::vector<float> h((size_t)B * N3);
    for (size_t i = 0; i < h.size(); ++i) h[i] = (float)std::sin(0.001 * (double)(i % 100003));
    NUMA_HIP_CHECK(hipMemcpy(r.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    NUMA_HIP_CHECK(hipMemcpy(z.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    auto run = [&]() {
        if (what == 0)
            apply(n, r.as<float>(), z.as<float>());
        else
            loss_grad(n, B, r.as<float>(), z.as<float>(), true, NUMA_NET_LOSS_RESIDUAL, nullptr);
    };
    run();
    hipEvent_t e0, e1;
    NUMA_HIP_CHECK(hipEventCreate(&e0));
    NUMA_HIP_CHECK(hipEventCreate(&e1));
    NUMA_HIP_CHECK(hipEventRecord(e0, c_.stream));
    for (int64_t i = 0; i < reps; ++i) run();
    NUMA_HIP_CHECK(hipEventRecord(e1, c_.stream));
    NUMA_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    NUMA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / (double)reps;
}

}  // namespace numa_hip

// ===========================================================================
// C entry points
// ===========================================================================
struct numa_hip_net {
    numa_hip::Ctx* ctx;
    std::shared_ptr<numa_hip::ConvNet> impl;
};

using numa_hip::require;

extern "C" {

int numa_hip_net_create(numa_hip_ctx* ctx, int64_t seed, numa_hip_net** out) {
    NUMA_API_BEGIN
    require(ctx && ctx->impl && out, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(ctx->impl->device));
    auto h = std::make_unique<numa_hip_net>();
    h->ctx = ctx->impl;
    h->impl = std::make_shared<numa_hip::ConvNet>(*ctx->impl, (uint64_t)seed);
    ++ctx->impl->live_handles;
    *out = h.release();
    NUMA_API_END
}

int numa_hip_net_destroy(numa_hip_net* net) {
    NUMA_API_BEGIN
    if (net) {
        NUMA_HIP_CHECK(hipStreamSynchronize(net->ctx->stream));
        --net->ctx->live_handles;
        delete net;
    }
    NUMA_API_END
}

int numa_hip_net_param_count(int64_t* count) {
    NUMA_API_BEGIN
    require(count != nullptr, "null argument");
    *count = numa_hip::cnn::kParams;
    NUMA_API_END
}

int numa_hip_net_get_params(numa_hip_net* net, float* host) {
    NUMA_API_BEGIN
    require(net && host, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(net->ctx->device));
    net->impl->get_params(host);
    NUMA_API_END
}

int numa_hip_net_set_params(numa_hip_net* net, const float* host) {
    NUMA_API_BEGIN
    require(net && host, "null argument");
    for (int64_t i = 0; i < numa_hip::cnn::kParams; ++i)
        require(std::isfinite(host[i]), "parameters must be finite");
    NUMA_HIP_CHECK(hipSetDevice(net->ctx->device));
    net->impl->set_params(host);
    NUMA_API_END
}

int numa_hip_net_forward(numa_hip_net* net, int64_t n, int64_t batch, const float* x_dev,
                         float* out_dev) {
    NUMA_API_BEGIN
    require(net && x_dev && out_dev, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(net->ctx->device));
    net->impl->forward(n, (int)std::min<int64_t>(batch, 1 << 20), x_dev, out_dev);
    NUMA_API_END
}

int numa_hip_net_loss_grad(numa_hip_net* net, int64_t n, int64_t batch, const float* x_dev,
                           const float* y_dev, int64_t loss_kind, double* loss_host,
                           float* grad_host) {
    NUMA_API_BEGIN
    require(net && x_dev && y_dev, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(net->ctx->device));
    auto& N = *net->impl;
    N.loss_grad(n, (int)std::min<int64_t>(batch, 1 << 20), x_dev, y_dev, false, loss_kind,
                N.mean_slot());
    if (loss_host) *loss_host = N.read_mean();
    if (grad_host) N.get_grads(grad_host);
    N.ctx().sync();
    NUMA_API_END
}

int numa_hip_net_train_step(numa_hip_net* net, int64_t n, int64_t batch, const float* r_dev,
                            const float* z_dev, int64_t loss_kind, double lr, double* loss_host) {
    NUMA_API_BEGIN
    require(net && r_dev && z_dev, "null argument");
    require(lr > 0.0, "learning rate must be positive");
    NUMA_HIP_CHECK(hipSetDevice(net->ctx->device));
    auto& N = *net->impl;
    N.loss_grad(n, (int)std::min<int64_t>(batch, 1 << 20), r_dev, z_dev, true, loss_kind,
                N.mean_slot());
    N.adam(lr);
    if (loss_host) *loss_host = N.read_mean();
    NUMA_API_END
}

int numa_hip_net_apply(numa_hip_net* net, int64_t n, const float* r_dev, float* z_dev) {
    NUMA_API_BEGIN
    require(net && r_dev && z_dev, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(net->ctx->device));
    net->impl->apply(n, r_dev, z_dev);
    NUMA_API_END
}

int numa_hip_net_train_solver(numa_hip_net* net, const numa_net_train_config* cfg,
                              double* loss_hist_host, int64_t hist_len,
                              numa_net_train_stats* stats) {
    NUMA_API_BEGIN
    require(net && cfg && stats, "null argument");
    require(hist_len >= 0, "negative history length");
    NUMA_HIP_CHECK(hipSetDevice(net->ctx->device));
    net->impl->train_from_solver(*cfg, loss_hist_host, hist_len, *stats);
    NUMA_API_END
}

int numa_hip_net_cost(numa_hip_net* net, int64_t n, int64_t what, int64_t batch, int64_t reps,
                      double* ms_per_call) {
    NUMA_API_BEGIN
    require(net && ms_per_call, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(net->ctx->device));
    *ms_per_call = net->impl->cost(n, what, batch, reps);
    NUMA_API_END
}

int numa_hip_solver_set_net(numa_hip_solver* s, numa_hip_net* net, double fallback_ratio) {
    NUMA_API_BEGIN
    require(s != nullptr, "null solver");
    require(!net || net->ctx == s->ctx, "network and solver belong to different contexts");
    s->impl->set_learned(net ? net->impl : nullptr, fallback_ratio);
    NUMA_API_END
}

}  // extern "C"
