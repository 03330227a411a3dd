// chandra.hip - Chandrasekhar H-equation on gfx950 for a batch of parameters
// c_b, solved by lockstep Newton-GMRES (methods/krylov_newton.hpp:69-114 with
// linalg/gmres.hpp:61-223 and no preconditioner, as the reference's
// apps/smoke_chandrasekhar.cpp and tests do).
//
// Operator (problems/chandrasekhar_op.hpp:31-136), per instance b:
//   (A x)_i = c_b sum_j K_ij x_j,  K_ij = 0.5 w_j mu_i / (mu_i + mu_j)
//   form 0 (reference):  F = 1 + x (A x) - x,   J v = v (Ax) + x (Av) - v
//   form 1 (Kelley):     F = x - 1/(1 - A x),   J v = v - (A v)/(1 - Ax)^2
// rule 0: Gauss-Legendre on [0,1] (functions/gauss_legendre_dynamic.hpp:50-114,
// computed on the host in long double); rule 1: midpoint mu_i = (i-1/2)/N,
// w = 1/N.
//
// All instances share one MFMA contraction per operator application
// (chandra_gemm.hpp): FP64 for F(x) always, and for J v either FP64
// (NUMA_PRECISION_FP64) or split-BF16 with FP32 accumulation
// (NUMA_PRECISION_MIXED).  Everything else (Krylov basis, Givens, update) is
// FP64.  Each instance follows the reference's step sequence exactly and keeps
// its own forcing term, Krylov counters, line search and convergence; a
// finished instance stops counting and stops changing its iterate.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <limits>
#include <memory>
#include <utility>
#include <vector>

#include "chandra_gemm.hpp"
#include "common.hpp"

namespace numa_hip {

// ---------------------------------------------------------------------------
// batched kernels over the padded [Bp][Np] layout; grid (x chunks, B)
// ---------------------------------------------------------------------------
constexpr int kChBX = 16;  // blocks per instance of the batched streaming kernels
// K is stored (FP64, plus bf16 hi / lo for split-BF16) when it takes at most
// this many bytes; NUMA_HIP_CHANDRA_STORE_K=0 in the environment at creation
// keeps the generating kernels instead.
constexpr size_t kChKStoreMax = size_t(16) << 30;
constexpr int kKdRT = 1, kKdCT = 4;  // k_gemm_f64_kd row, column tiles per wave
constexpr int kKdNarrow = -1;        // packed FP64 block: column tile by width (gemm_f64)
// k_gemm_bf16x3_k: waves, row tiles per wave; 4 column tiles per wave (k chunk
// 64) where the padded batch is a multiple of 128, else 2 (k chunk 128)
constexpr int kKbWR = 4, kKbRT = 1;
static_assert(kChPadN % (64 * kKdRT) == 0 && kChPadN % (kKbWR * 32 * kKbRT) == 0 &&
                  kChPadN % (4 * kGbK) == 0 && kChPadB % (16 * kKdCT) == 0 &&
                  kChPadB % 64 == 0,
              "padded N and batch must hold whole contraction tiles and split-K chunks");

struct BEpi {   // per-instance epilogue of a batched reduction
    int kind;   // EPI_STORE / EPI_SQRT / EPI_NORM_SCALE
    const double* mul;  // EPI_STORE multiplier per instance (null = 1)
    double* out;
    double* out2;
};

// Kernels over a selection of instances take `imap` (null: every instance):
// workgroup row y serves instance imap[y].
__global__ __launch_bounds__(256) void k_bfinish(int nbx, const double* partials, BEpi ep,
                                                  const int* imap) {
    const int b = imap ? imap[blockIdx.x] : (int)blockIdx.x;
    double s = 0.0;
    for (int x = threadIdx.x; x < nbx; x += 256) s += partials[b * nbx + x];
    s = block_sum<256>(s);
    if (threadIdx.x == 0) {
        if (ep.kind == EPI_STORE) {
            ep.out[b] = s * (ep.mul ? ep.mul[b] : 1.0);
        } else {
            const double r = sqrt(s);
            ep.out[b] = r;
            if (ep.kind == EPI_NORM_SCALE && ep.out2) ep.out2[b] = r > 0.0 ? 1.0 / r : 0.0;
        }
    }
}

__global__ __launch_bounds__(256) void k_bdot(int N, int Np, const double* a, const double* c,
                                               Reduce red, const int* imap) {
    const int b = imap ? imap[blockIdx.y] : (int)blockIdx.y;
    const double* pa = a + (size_t)b * Np;
    const double* pc = c + (size_t)b * Np;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) acc += pa[i] * pc[i];
    block_partial<256>(acc, red, b * gridDim.x + blockIdx.x);
}

// w_b -= (h_b s_b) V_b ;  partial of <w_b, next_b> or <w_b, w_b> (next == null)
__global__ __launch_bounds__(256) void k_bmgs(int N, int Np, double* w, const double* V,
                                               const double* h, const double* s,
                                               const double* next, Reduce red, const int* imap) {
#pragma clang fp contract(off)
    const int b = imap ? imap[blockIdx.y] : (int)blockIdx.y;
    const double coef = h[b] * s[b];
    double* pw = w + (size_t)b * Np;
    const double* pv = V + (size_t)b * Np;
    const double* pn = next ? next + (size_t)b * Np : nullptr;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        const double nv = pn ? pn[i] : 0.0;
        const double wn = pw[i] - coef * pv[i];
        pw[i] = wn;
        acc += pn ? wn * nv : wn * wn;
    }
    block_partial<256>(acc, red, b * gridDim.x + blockIdx.x);
}

// From the contraction Y = K X of the point X: Ax = c_b Y, F(X) and ||F||^2.
__global__ __launch_bounds__(256) void k_ch_F(int N, int Np, int form, const double* c,
                                               const double* X, const double* Y, double* Ax,
                                               double* F, Reduce red) {
#pragma clang fp contract(off)
    const int b = blockIdx.y;
    const size_t o = (size_t)b * Np;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        const double x = X[o + i];
        const double ax = c[b] * Y[o + i];
        const double f = form == 0 ? 1.0 + x * ax - x : x - 1.0 / (1.0 - ax);
        Ax[o + i] = ax;
        F[o + i] = f;
        acc += f * f;
    }
    block_partial<256>(acc, red, b * gridDim.x + blockIdx.x);
}

// out = scale_b * J(X) V from Y = K V (Av = c_b Y); optionally out = r - J V
// (restart residual, then the partial is ||out||^2).
// Y is the FP64 contraction, or (P non-null) the ksplit FP32 split-K slices of
// the split-BF16 contraction, slice stride `slice`, summed here in the order
// k_splitk_reduce uses (so both give the same Y).
__global__ __launch_bounds__(256) void k_ch_jv(int N, int Np, int form, const double* c,
                                                const double* X, const double* Ax,
                                                const double* V, const double* Y,
                                                const float* P, int ksplit, int64_t slice,
                                                const double* scale, const double* rhs,
                                                double* out, Reduce red, const int* imap,
                                                const double* dotv) {
#pragma clang fp contract(off)
    const int b = imap ? imap[blockIdx.y] : (int)blockIdx.y;
    const size_t o = (size_t)b * Np;
    const double sc = scale ? scale[b] : 1.0;
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        double y;
        if (P) {
            y = 0.0;
            for (int k = 0; k < ksplit; ++k) y += (double)P[(size_t)k * slice + o + i];
        } else {
            y = Y[o + i];
        }
        const double v = V[o + i], av = c[b] * y, ax = Ax[o + i];
        double jv;
        if (form == 0) {
            jv = v * ax + X[o + i] * av - v;
        } else {
            const double q = 1.0 - ax;
            jv = v - av / (q * q);
        }
        jv = jv * sc;
        if (rhs) jv = rhs[o + i] - jv;
        out[o + i] = jv;
        if (dotv)  // <out, dotv>, term for term as k_bdot adds it (one fused multiply-add)
            acc = __builtin_fma(jv, dotv[o + i], acc);
        else
            acc += jv * jv;
    }
    block_partial<256>(acc, red, b * gridDim.x + blockIdx.x);
}

// out = F''(X)[V, V] from Y = K V (Av = c_b Y): 2 v Av (form 0,
// chandrasekhar_op.hpp:81-93) or -2 (Av)^2 / q^3 with q = 1 - Ax (form 1).
__global__ __launch_bounds__(256) void k_ch_d2(int N, int Np, int form, const double* c,
                                                const double* Ax, const double* V,
                                                const double* Y, double* out) {
#pragma clang fp contract(off)
    const int b = blockIdx.y;
    const size_t o = (size_t)b * Np;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        const double av = c[b] * Y[o + i];
        double d;
        if (form == 0) {
            d = 2.0 * V[o + i] * av;
        } else {
            const double q = 1.0 - Ax[o + i];
            d = -2.0 * (av * av) / (q * q * q);
        }
        out[o + i] = d;
    }
}

// Xt_b = (X_b - Z_b) - 0.5 U_b: the Chebyshev step (krylov_newton.hpp:162)
__global__ __launch_bounds__(256) void k_ch_cheb_trial(int N, int Np, const double* X,
                                                        const double* Z, const double* U,
                                                        double* Xt) {
#pragma clang fp contract(off)
    const int b = blockIdx.y;
    const size_t o = (size_t)b * Np;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
        Xt[o + i] = X[o + i] - Z[o + i] - 0.5 * U[o + i];
}

// Xt_b = X_b - lam_b Z_b
__global__ __launch_bounds__(256) void k_ch_trial(int N, int Np, const double* X, const double* Z,
                                                   const double* lam, double* Xt) {
#pragma clang fp contract(off)
    const int b = blockIdx.y;
    const size_t o = (size_t)b * Np;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
        Xt[o + i] = X[o + i] - lam[b] * Z[o + i];
}

// dst_b = src_b where flag_b != 0
__global__ __launch_bounds__(256) void k_ch_select(int N, int Np, const double* flag,
                                                    const double* src, double* dst) {
    const int b = blockIdx.y;
    if (flag[b] == 0.0) return;
    const size_t o = (size_t)b * Np;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) dst[o + i] = src[o + i];
}

// y_b += sum_k coef[k*B + b] V_k,b   (k < count[b]: the basis vectors built
// while instance b was iterating; later ones may hold another instance's
// leftovers, see the packed contractions)
__global__ __launch_bounds__(256) void k_ch_combine(int N, int Np, int B, const int* count,
                                                     const double* const* V, const double* coef,
                                                     double* y) {
#pragma clang fp contract(off)
    const int b = blockIdx.y;
    const int kb = count[b];
    if (kb == 0) return;
    const size_t o = (size_t)b * Np;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        double s = y[o + i];
        for (int k = 0; k < kb; ++k) s += coef[k * B + b] * V[k][o + i];
        y[o + i] = s;
    }
}

// Packed contractions: when only some instances use a contraction's result
// (the rest have converged in the lockstep iteration), their columns are
// gathered into a packed block of npad columns (a multiple of the
// contraction's column tile), contracted, and the results scattered back.
// Every column of a contraction is computed by the same sequence of
// operations wherever it sits in the block, so a packed result is bit for
// bit the full one.  dst column p = src column map[p] (p < nact), 0 for
// nact <= p < npad.
__global__ __launch_bounds__(256) void k_ch_gather(int Np, int nact, const int* map,
                                                    const double* src, double* dst) {
    const int p = blockIdx.y;
    const bool in = p < nact;
    const size_t so = in ? (size_t)map[p] * Np : 0, o = (size_t)p * Np;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < Np; i += gridDim.x * 256)
        dst[o + i] = in ? src[so + i] : 0.0;
}
// dst column map[p] = src column p (p < nact = gridDim.y), for each of
// gridDim.z slices (strides sstride / dstride elements).
template <class T>
__global__ __launch_bounds__(256) void k_ch_scatter(int Np, const int* map, const T* src, T* dst,
                                                     int64_t sstride, int64_t dstride) {
    const int p = blockIdx.y;
    const T* in = src + (int64_t)blockIdx.z * sstride + (int64_t)p * Np;
    T* out = dst + (int64_t)blockIdx.z * dstride + (int64_t)map[p] * Np;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < Np; i += gridDim.x * 256) out[i] = in[i];
}

// user [B][N] <-> padded [Bp][Np]
__global__ __launch_bounds__(256) void k_ch_pack(int N, int Np, const double* src, double* dst,
                                                  int to_padded) {
    const int b = blockIdx.y;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        if (to_padded)
            dst[(size_t)b * Np + i] = src[(size_t)b * N + i];
        else
            dst[(size_t)b * N + i] = src[(size_t)b * Np + i];
    }
}

// ---------------------------------------------------------------------------
// quadrature rules (host)
// ---------------------------------------------------------------------------
// gauss_legendre_unit (functions/gauss_legendre_dynamic.hpp:50-114): Newton on
// the Legendre recurrence in long double, weights halved to sum to 1.
inline void gauss_legendre_unit(size_t n, std::vector<double>& nodes, std::vector<double>& weights) {
    nodes.assign(n, 0.0);
    weights.assign(n, 0.0);
    const size_t half = (n + 1) / 2;
    const long double pi = std::acos(-1.0L);
    const long double eps = 32.0L * std::numeric_limits<long double>::epsilon();
    auto leg = [n](long double x, long double& p, long double& dp) {
        long double p0 = 1.0L, p1 = x;
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
        bool ok = false;
        for (int it = 0; it < 100; ++it) {
            long double p, dp;
            leg(z, p, dp);
            const long double d = p / dp;
            z -= d;
            if (std::abs(d) <= eps * (1.0L + std::abs(z))) {
                ok = true;
                break;
            }
        }
        require(ok, "Gauss-Legendre Newton iteration did not converge", NUMA_ERR_HIP);
        long double p, dp;
        leg(z, p, dp);
        const long double wpm = 2.0L / ((1.0L - z * z) * dp * dp);
        nodes[i] = (double)(0.5L * (1.0L - z));
        nodes[n - 1 - i] = (double)(0.5L * (1.0L + z));
        weights[i] = (double)(0.5L * wpm);
        weights[n - 1 - i] = (double)(0.5L * wpm);
    }
}

// ---------------------------------------------------------------------------
// batched operator + solver
// ---------------------------------------------------------------------------
class ChandraBatch {
public:
    ChandraBatch(Ctx& c, int64_t n, int64_t B, const double* cvals, int rule, int form,
                 int precision, int64_t restart)
        : c_(c), N_(n), B_(B), form_(form), prec_(precision), m_(restart) {
        require(n > 0 && B > 0, "Chandrasekhar needs N > 0 and batch > 0");
        require(n <= (1 << 20) && B * kChBX <= kMaxPartials, "Chandrasekhar size too large",
                NUMA_ERR_OVERFLOW);
        require(rule == 0 || rule == 1, "rule must be 0 (Gauss-Legendre) or 1 (midpoint)");
        require(form == 0 || form == 1, "form must be 0 (reference) or 1 (Kelley)");
        require(precision == NUMA_PRECISION_FP64 || precision == NUMA_PRECISION_MIXED,
                "unknown precision");
        require(restart >= 1 && restart <= 256, "gmres restart must be in [1, 256]");
        for (int64_t b = 0; b < B; ++b)
            require(std::isfinite(cvals[b]) && cvals[b] >= 0.0 && cvals[b] <= 1.0,
                    "c must be finite in [0, 1] (chandrasekhar_op.hpp:40-43)");
        Np_ = (n + kChPadN - 1) / kChPadN * kChPadN;
        Bp_ = (B + kChPadB - 1) / kChPadB * kChPadB;
        // quadrature, padded with w = 0, mu = 1
        std::vector<double> mu, w;
        if (rule == 0) {
            gauss_legendre_unit((size_t)n, mu, w);
        } else {
            mu.resize(n);
            w.assign(n, 1.0 / double(n));
            for (int64_t i = 0; i < n; ++i) mu[i] = (double(i) + 0.5) / double(n);
        }
        mu.resize(Np_, 1.0);
        w.resize(Np_, 0.0);
        std::vector<float> muf(mu.begin(), mu.end()), wf(w.begin(), w.end());
        std::vector<double> cv(cvals, cvals + B);
        cv.resize(Bp_, 0.0);
        mu_ = upload(mu);
        w_ = upload(w);
        c_dev_ = upload(cv);
        muf_ = static_cast<float*>(alloc(Np_ * sizeof(float)));
        wf_ = static_cast<float*>(alloc(Np_ * sizeof(float)));
        NUMA_HIP_CHECK(hipMemcpy(muf_, muf.data(), Np_ * sizeof(float), hipMemcpyHostToDevice));
        NUMA_HIP_CHECK(hipMemcpy(wf_, wf.data(), Np_ * sizeof(float), hipMemcpyHostToDevice));
        const size_t vec = (size_t)Bp_ * Np_;
        for (double** p : {&X_, &Xt_, &F_, &Ft_, &Ax_, &Axt_, &Z_, &Y_, &R_})
            *p = static_cast<double*>(zalloc(vec * sizeof(double)));
        V_.resize((size_t)m_ + 1);
        for (auto& v : V_) v = static_cast<double*>(zalloc(vec * sizeof(double)));
        if (prec_ == NUMA_PRECISION_MIXED) {
            Xh_ = static_cast<__bf16*>(zalloc(vec * sizeof(__bf16)));
            Xl_ = static_cast<__bf16*>(zalloc(vec * sizeof(__bf16)));
            // split K until the stored-K grid has 512 workgroups (at most 4 slices);
            // the generating kernel uses the same slices, so both return the same bits
            kb_ct_ = Bp_ % 128 == 0 ? 4 : 2;
            const int64_t tiles = (Np_ / (kKbWR * 32 * kKbRT)) * (Bp_ / (32 * kb_ct_));
            ksplit_ = 1;
            while (ksplit_ < 4 && tiles * ksplit_ < 512) ksplit_ *= 2;
            P_ = static_cast<float*>(zalloc((size_t)ksplit_ * vec * sizeof(float)));
        }
        // per-instance scalars: H ((m+1) x m x B), s ((m+1) x B), misc
        nsc_ = ((m_ + 1) * m_ + (m_ + 1) + 8) * Bp_;
        sd_ = static_cast<double*>(zalloc((size_t)nsc_ * sizeof(double)));
        void* hp = nullptr;
        NUMA_HIP_CHECK(hipHostMalloc(&hp, (size_t)nsc_ * sizeof(double), hipHostMallocDefault));
        hs_ = static_cast<double*>(hp);
        vptr_ = static_cast<const double**>(alloc((size_t)(m_ + 1) * sizeof(double*)));
        coef_ = static_cast<double*>(alloc((size_t)(m_ + 1) * Bp_ * sizeof(double)));
        NUMA_HIP_CHECK(hipMemcpy(vptr_, V_.data(), (m_ + 1) * sizeof(double*), hipMemcpyHostToDevice));
        kcnt_ = static_cast<int*>(zalloc((size_t)Bp_ * sizeof(int)));
        gmap_ = static_cast<int*>(zalloc((size_t)Bp_ * sizeof(int)));
        void* hm = nullptr;
        NUMA_HIP_CHECK(hipHostMalloc(&hm, (size_t)Bp_ * sizeof(int), hipHostMallocDefault));
        gmap_h_ = static_cast<int*>(hm);
        store_k();
        const char* pk = std::getenv("NUMA_HIP_CHANDRA_PACK");
        lean_ = !(pk && pk[0] == '0' && pk[1] == 0);
        if (Kd_ && lean_) {
            pack_ = true;
            Xg_ = static_cast<double*>(zalloc(vec * sizeof(double)));
            Yg_ = static_cast<double*>(zalloc(vec * sizeof(double)));
            if (Kh_) Pg_ = static_cast<float*>(zalloc((size_t)ksplit_ * vec * sizeof(float)));
        }
    }
    ~ChandraBatch() {
        for (void* p : allocs_) (void)hipFree(p);
        if (hs_) (void)hipHostFree(hs_);
        if (gmap_h_) (void)hipHostFree(gmap_h_);
    }

    int64_t size() const { return N_ * B_; }

    // F for all instances at x (user layout [B][N]); ||F_b|| to norms (host).
    void eval_user(const double* x, double* F, double* norms) {
        pack(x, X_, true);
        eval_F(X_, F_, Ax_, MISC(0));
        pack(F_, F, false);
        if (norms) read(MISC(0), B_, norms);
        c_.sync();
    }

    // out = J(x) v for all instances (user layout), precision of the handle.
    void jvp_user(const double* x, const double* v, double* out) {
        pack(x, X_, true);
        eval_F(X_, F_, Ax_, MISC(0));
        pack(v, V_[0], true);
        jv(V_[0], nullptr, nullptr, R_, nullptr);
        pack(R_, out, false);
        c_.sync();
    }

    // out = F''(x)[v, v] for all instances (user layout), FP64.
    void d2_user(const double* x, const double* v, double* out) {
        pack(x, X_, true);
        eval_F(X_, F_, Ax_, MISC(0));
        pack(v, V_[0], true);
        d2(V_[0], R_);
        pack(R_, out, false);
        c_.sync();
    }

    // NUMA_METHOD_NEWTON (newton_krylov) or NUMA_METHOD_CHEBYSHEV
    // (chebyshev_inexact) for solve().
    void set_method(int method) {
        require(method == NUMA_METHOD_NEWTON || method == NUMA_METHOD_CHEBYSHEV,
                "the Chandrasekhar solver runs Newton or inexact Chebyshev");
        method_ = method;
        if (method == NUMA_METHOD_CHEBYSHEV && !W_) {
            W_ = static_cast<double*>(zalloc((size_t)Bp_ * Np_ * sizeof(double)));
            U_ = static_cast<double*>(zalloc((size_t)Bp_ * Np_ * sizeof(double)));
        }
    }

    // Average ms of one batched contraction and its useful FLOP count 2 N^2 B.
    double gemm_cost(int precision, int64_t reps, double* flops) {
        require(reps > 0, "reps must be positive");
        require(precision == NUMA_PRECISION_FP64 || prec_ == NUMA_PRECISION_MIXED,
                "the handle was created without the split-BF16 buffers");
        launch_fill_rows(V_[0], 0.5);
        auto run = [&] {
            if (precision == NUMA_PRECISION_FP64)
                gemm_f64(V_[0], Y_);
            else
                gemm_bf16(V_[0], Y_);
        };
        run();
        hipEvent_t e0, e1;
        NUMA_HIP_CHECK(hipEventCreate(&e0));
        NUMA_HIP_CHECK(hipEventCreate(&e1));
        NUMA_HIP_CHECK(hipEventRecord(e0, c_.stream));
        for (int64_t r = 0; r < reps; ++r) run();
        NUMA_HIP_CHECK(hipEventRecord(e1, c_.stream));
        NUMA_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        NUMA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (flops) *flops = 2.0 * (double)N_ * (double)N_ * (double)B_;
        return (double)ms / (double)reps;
    }

    void solve(const numa_krylov_params& p, double* xuser, numa_krylov_result* out);

private:
    double* MISC(int k) { return sd_ + ((m_ + 1) * m_ + (m_ + 1) + k) * Bp_; }
    double* H(int64_t i, int64_t j) { return sd_ + (j * (m_ + 1) + i) * Bp_; }
    double* S(int64_t i) { return sd_ + ((m_ + 1) * m_ + i) * Bp_; }

    // Stores K after every buffer the handle needs, and only if it fits in
    // kChKStoreMax and the device has the memory: when an allocation fails the
    // handle keeps the generating kernels (same results).
    void store_k() {
        const char* sk = std::getenv("NUMA_HIP_CHANDRA_STORE_K");
        const size_t kbytes =
            (size_t)Np_ * (ldd() * 8 + (prec_ == NUMA_PRECISION_MIXED ? ldb() * 4 : 0));
        if ((sk && sk[0] == '0' && sk[1] == 0) || kbytes > kChKStoreMax) return;
        void* kd = nullptr;
        void* kh = nullptr;
        void* kl = nullptr;
        bool ok = hipMalloc(&kd, (size_t)Np_ * ldd() * sizeof(double)) == hipSuccess;
        if (ok && prec_ == NUMA_PRECISION_MIXED)
            ok = hipMalloc(&kh, (size_t)Np_ * ldb() * sizeof(__bf16)) == hipSuccess &&
                 hipMalloc(&kl, (size_t)Np_ * ldb() * sizeof(__bf16)) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();  // the out-of-memory error is handled here
            for (void* q : {kd, kh, kl})
                if (q) (void)hipFree(q);
            return;
        }
        for (void* q : {kd, kh, kl})
            if (q) allocs_.push_back(q);
        Kd_ = static_cast<double*>(kd);
        Kh_ = static_cast<__bf16*>(kh);
        Kl_ = static_cast<__bf16*>(kl);
        hipLaunchKernelGGL(k_fill_k, dim3((unsigned)((Np_ + 255) / 256), (unsigned)Np_), dim3(256),
                           0, c_.stream, (int)Np_, (int)ldd(), (int)ldb(), (const double*)mu_,
                           (const double*)w_, (const float*)muf_, (const float*)wf_, Kd_, Kh_,
                           Kl_);
        NUMA_HIP_LAUNCH_CHECK();
        c_.sync();
    }

    void* alloc(size_t bytes) {
        void* p = nullptr;
        NUMA_HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 8)));
        allocs_.push_back(p);
        return p;
    }
    void* zalloc(size_t bytes) {
        void* p = alloc(bytes);
        NUMA_HIP_CHECK(hipMemset(p, 0, bytes));
        return p;
    }
    double* upload(const std::vector<double>& v) {
        double* p = static_cast<double*>(alloc(v.size() * sizeof(double)));
        NUMA_HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        return p;
    }
    void read(const double* dev, int64_t count, double* host) {
        NUMA_HIP_CHECK(hipMemcpyAsync(hs_, dev, (size_t)count * sizeof(double),
                                      hipMemcpyDeviceToHost, c_.stream));
        c_.sync();
        std::copy(hs_, hs_ + count, host);
    }
    void write(double* dev, const std::vector<double>& host) {
        std::copy(host.begin(), host.end(), hs_);
        NUMA_HIP_CHECK(hipMemcpyAsync(dev, hs_, host.size() * sizeof(double),
                                      hipMemcpyHostToDevice, c_.stream));
        c_.sync();
    }
    dim3 bgrid() const { return dim3((unsigned)std::min<int64_t>(kChBX, (N_ + 255) / 256), (unsigned)B_); }
    unsigned nbx() const { return bgrid().x; }

    void pack(const double* src, double* dst, bool to_padded) {
        hipLaunchKernelGGL(k_ch_pack, bgrid(), dim3(256), 0, c_.stream, (int)N_, (int)Np_, src, dst,
                           to_padded ? 1 : 0);
        NUMA_HIP_LAUNCH_CHECK();
    }
    void launch_fill_rows(double* dst, double v) {
        std::vector<double> row((size_t)N_, v);
        for (int64_t b = 0; b < B_; ++b)
            NUMA_HIP_CHECK(hipMemcpyAsync(dst + b * Np_, row.data(), N_ * sizeof(double),
                                          hipMemcpyHostToDevice, c_.stream));
        c_.sync();
    }
    void bfinish(const BEpi& ep) {
        if (nrows() == 0) return;
        hipLaunchKernelGGL(k_bfinish, dim3(nrows()), dim3(256), 0, c_.stream, (int)nbx(),
                           (const double*)c_.red.partials, ep, imap());
        NUMA_HIP_LAUNCH_CHECK();
    }

    // Columns of a packed contraction for the instances sel[b] != 0: uploads
    // their map and returns the packed width (nact rounded up to `tile`), or 0
    // when packing would not drop a column tile (then the full contraction
    // runs).
    int64_t pack_cols(const std::vector<char>* sel, int64_t tile, int64_t& nact) {
        if (!pack_ || !sel) return 0;
        const std::vector<int> map = instance_list(*sel);
        nact = (int64_t)map.size();
        if (tile == kKdNarrow) tile = nact <= 32 ? 16 : nact <= 64 ? 32 : 16 * kKdCT;
        const int64_t npad = std::max<int64_t>(1, (nact + tile - 1) / tile) * tile;
        if (npad >= Bp_) return 0;
        upload_map(map);
        return npad;
    }
    std::vector<int> instance_list(const std::vector<char>& sel) const {
        std::vector<int> map;
        for (int64_t b = 0; b < B_; ++b)
            if (sel[(size_t)b]) map.push_back((int)b);
        return map;
    }
    // the instance list in gmap_ (uploaded when it changed)
    void upload_map(const std::vector<int>& map) {
        if (map.empty() || map == gmap_last_) return;
        c_.sync();  // the pinned staging may still feed the last upload
        std::copy(map.begin(), map.end(), gmap_h_);
        NUMA_HIP_CHECK(hipMemcpyAsync(gmap_, gmap_h_, map.size() * sizeof(int),
                                      hipMemcpyHostToDevice, c_.stream));
        gmap_last_ = map;
    }
    // The instances the per-instance kernels (J v, dots, MGS, finishes) run
    // on until the next call: sel[b] != 0, or every instance (null).  The
    // others' outputs are left as they were.
    void select_instances(const std::vector<char>* sel) {
        if (!sel || !lean_) {
            nsel_ = -1;
            return;
        }
        const std::vector<int> map = instance_list(*sel);
        if ((int64_t)map.size() == B_) {
            nsel_ = -1;
            return;
        }
        upload_map(map);
        nsel_ = (int64_t)map.size();
    }
    const int* imap() const { return nsel_ < 0 ? nullptr : gmap_; }
    unsigned nrows() const { return (unsigned)(nsel_ < 0 ? B_ : nsel_); }
    dim3 sgrid() const { return dim3(nbx(), nrows()); }
    void gather(const double* src, int64_t nact, int64_t npad) {
        hipLaunchKernelGGL(k_ch_gather, dim3((unsigned)nbx(), (unsigned)npad), dim3(256), 0,
                           c_.stream, (int)Np_, (int)nact, (const int*)gmap_, src, Xg_);
        NUMA_HIP_LAUNCH_CHECK();
    }

    // Yout = K Xin (columns sel only, when given)
    void gemm_f64(const double* Xin, double* Yout, const std::vector<char>* sel = nullptr) {
        int64_t nact = 0;
        // up to 128 columns: 4 chunks of loads in flight, with narrower wave
        // tiles (16 / 32 / 64 columns) for narrower blocks, which stream all
        // of K for few MFMAs; past 128 the full-width kernel.  Each element's
        // chain of MFMA steps is the same in every form.
        const int64_t npad = Kd_ ? pack_cols(sel, kKdNarrow, nact) : 0;
        if (npad > 0) {
            if (nact == 0) return;
            gather(Xin, nact, npad);
            const unsigned gx = (unsigned)(Np_ / (64 * kKdRT));
            static_assert(kChPadN % (kGdK * 4) == 0, "prefetch ring of 4 chunks");
            auto pf = [&](auto kern, int ct) {
                hipLaunchKernelGGL(kern, dim3(gx, (unsigned)(npad / (16 * ct))), dim3(256), 0,
                                   c_.stream, (int)Np_, (int)ldd(), (const double*)Kd_,
                                   (const double*)Xg_, Yg_);
            };
            if (npad <= 32)
                pf(k_gemm_f64_kd_pf<kKdRT, 1, 4>, 1);
            else if (npad <= 64)
                pf(k_gemm_f64_kd_pf<kKdRT, 2, 4>, 2);
            else if (npad <= 128)
                pf(k_gemm_f64_kd_pf<kKdRT, 4, 4>, 4);
            else
                hipLaunchKernelGGL((k_gemm_f64_kd<kKdRT, kKdCT>),
                                   dim3(gx, (unsigned)(npad / (16 * kKdCT))), dim3(256), 0,
                                   c_.stream, (int)Np_, (int)ldd(), (const double*)Kd_,
                                   (const double*)Xg_, Yg_);
            NUMA_HIP_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_ch_scatter<double>, dim3((unsigned)nbx(), (unsigned)nact, 1),
                               dim3(256), 0, c_.stream, (int)Np_, (const int*)gmap_,
                               (const double*)Yg_, Yout, (int64_t)0, (int64_t)0);
            NUMA_HIP_LAUNCH_CHECK();
            return;
        }
        if (Kd_) {
            hipLaunchKernelGGL((k_gemm_f64_kd<kKdRT, kKdCT>),
                               dim3((unsigned)(Np_ / (64 * kKdRT)), (unsigned)(Bp_ / (16 * kKdCT))),
                               dim3(256), 0, c_.stream, (int)Np_, (int)ldd(), (const double*)Kd_,
                               Xin, Yout);
            NUMA_HIP_LAUNCH_CHECK();
            return;
        }
        hipLaunchKernelGGL(k_gemm_f64, dim3((unsigned)(Np_ / kG64Rows), (unsigned)(Bp_ / kG64Cols)),
                           dim3(256), 0, c_.stream, (int)Np_, (const double*)mu_, (const double*)w_,
                           Xin, Yout);
        NUMA_HIP_LAUNCH_CHECK();
    }
    // Split X, contract into the split-K slices P_, and (Yout non-null) sum
    // them into Yout; the J v path passes null and sums them in k_ch_jv.
    // With sel, only the split-K slices of the selected columns are written
    // (to P_; Yout must be null).
    void gemm_bf16(const double* Xin, double* Yout, const std::vector<char>* sel = nullptr) {
        int64_t nact = 0;
        // a packed block uses 2 column tiles per wave (64 columns a workgroup)
        const int64_t npad = Kh_ && !Yout ? pack_cols(sel, 64, nact) : 0;
        const int ct = npad > 0 ? 2 : kb_ct_;
        if (npad > 0 && nact == 0) return;
        const int64_t cols = npad > 0 ? npad : Bp_;
        if (npad > 0) gather(Xin, nact, npad);
        const int64_t total = cols * Np_;
        hipLaunchKernelGGL(k_split_x, dim3(2048), dim3(256), 0, c_.stream, total,
                           npad > 0 ? (const double*)Xg_ : Xin, Xh_, Xl_);
        NUMA_HIP_LAUNCH_CHECK();
        if (Kh_) {
            const dim3 grid((unsigned)(Np_ / (kKbWR * 32 * kKbRT)), (unsigned)(cols / (32 * ct)),
                            (unsigned)ksplit_);
            const int64_t klen = Np_ / ksplit_;
            auto kern = ct == 4 ? (klen % 64 == 0 ? k_gemm_bf16x3_k<kKbWR, kKbRT, 4, 64>
                                                      : k_gemm_bf16x3_k<kKbWR, kKbRT, 4, 32>)
                                    : (klen % 128 == 0  ? k_gemm_bf16x3_k<kKbWR, kKbRT, 2, 128>
                                       : klen % 64 == 0 ? k_gemm_bf16x3_k<kKbWR, kKbRT, 2, 64>
                                                        : k_gemm_bf16x3_k<kKbWR, kKbRT, 2, 32>);
            hipLaunchKernelGGL(kern, grid, dim3(kKbWR * 64), 0, c_.stream, (int)Np_, (int)ldb(),
                               ksplit_,
                               (const __bf16*)Kh_, (const __bf16*)Kl_, (const __bf16*)Xh_,
                               (const __bf16*)Xl_, npad > 0 ? Pg_ : P_);
            if (npad > 0) {
                NUMA_HIP_LAUNCH_CHECK();
                hipLaunchKernelGGL(k_ch_scatter<float>,
                                   dim3((unsigned)nbx(), (unsigned)nact, (unsigned)ksplit_),
                                   dim3(256), 0, c_.stream, (int)Np_, (const int*)gmap_,
                                   (const float*)Pg_, P_, (int64_t)(npad * Np_),
                                   (int64_t)(Bp_ * Np_));
                NUMA_HIP_LAUNCH_CHECK();
                return;
            }
        } else {
            hipLaunchKernelGGL(k_gemm_bf16x3,
                               dim3((unsigned)(Np_ / kGbRows), (unsigned)(Bp_ / kGbCols),
                                    (unsigned)ksplit_),
                               dim3(256), 0, c_.stream, (int)Np_, ksplit_, (const float*)muf_,
                               (const float*)wf_, (const __bf16*)Xh_, (const __bf16*)Xl_, P_);
        }
        NUMA_HIP_LAUNCH_CHECK();
        if (!Yout) return;
        hipLaunchKernelGGL(k_splitk_reduce, dim3(2048), dim3(256), 0, c_.stream, total, ksplit_,
                           (const float*)P_, Yout);
        NUMA_HIP_LAUNCH_CHECK();
    }

    // F and Ax at X; ||F_b|| -> norms (sel: only these instances' values are
    // used; the others' are left meaningless)
    void eval_F(const double* X, double* F, double* Ax, double* norms,
                const std::vector<char>* sel = nullptr) {
        gemm_f64(X, Y_, sel);
        hipLaunchKernelGGL(k_ch_F, bgrid(), dim3(256), 0, c_.stream, (int)N_, (int)Np_, form_,
                           (const double*)c_dev_, X, (const double*)Y_, Ax, F, c_.red);
        NUMA_HIP_LAUNCH_CHECK();
        bfinish(BEpi{EPI_SQRT, nullptr, norms, nullptr});
    }

    // out = F''(X_)[V, V], FP64 contraction (Ax_ holds A X_)
    void d2(const double* V, double* out, const std::vector<char>* sel = nullptr) {
        gemm_f64(V, Y_, sel);
        hipLaunchKernelGGL(k_ch_d2, bgrid(), dim3(256), 0, c_.stream, (int)N_, (int)Np_, form_,
                           (const double*)c_dev_, (const double*)Ax_, V, (const double*)Y_, out);
        NUMA_HIP_LAUNCH_CHECK();
    }

    // out = scale * J(X_) V  (or rhs - J V); ||out_b||^2 partials -> ep if given
    // dotv: the partials are of <out_b, dotv_b> instead of ||out_b||^2
    void jv(const double* V, const double* scale, const double* rhs, double* out, const BEpi* ep,
            const std::vector<char>* sel = nullptr, const double* dotv = nullptr) {
        if (nrows() == 0) return;
        const bool split = prec_ == NUMA_PRECISION_MIXED;
        if (split)
            gemm_bf16(V, nullptr, sel);
        else
            gemm_f64(V, Y_, sel);
        hipLaunchKernelGGL(k_ch_jv, sgrid(), dim3(256), 0, c_.stream, (int)N_, (int)Np_, form_,
                           (const double*)c_dev_, (const double*)X_, (const double*)Ax_, V,
                           (const double*)Y_, split ? (const float*)P_ : nullptr, ksplit_,
                           (int64_t)Bp_ * Np_, scale, rhs, out, c_.red, imap(), dotv);
        NUMA_HIP_LAUNCH_CHECK();
        if (ep) bfinish(*ep);
    }

    void bdot(const double* a, const double* b, const BEpi& ep) {
        if (nrows() == 0) return;
        hipLaunchKernelGGL(k_bdot, sgrid(), dim3(256), 0, c_.stream, (int)N_, (int)Np_, a, b, c_.red,
                           imap());
        NUMA_HIP_LAUNCH_CHECK();
        bfinish(ep);
    }

    struct GState {
        bool active = false;  // participates in this GMRES
        numa_gmres_result res{};
    };
    void gmres_batch(const std::vector<char>& active, const std::vector<double>& rtol,
                     int64_t restart, int64_t max_iters, std::vector<numa_gmres_result>& out,
                     const double* rhs, double* sol);
    int method_ = NUMA_METHOD_NEWTON;
    double *W_ = nullptr, *U_ = nullptr;  // F''[z,z] and the second solve (Chebyshev)

    Ctx& c_;
    int64_t N_, B_, Np_ = 0, Bp_ = 0;
    int form_, prec_;
    int64_t m_;
    double *mu_ = nullptr, *w_ = nullptr, *c_dev_ = nullptr;
    float *muf_ = nullptr, *wf_ = nullptr;
    double *X_ = nullptr, *Xt_ = nullptr, *F_ = nullptr, *Ft_ = nullptr, *Ax_ = nullptr,
           *Axt_ = nullptr, *Z_ = nullptr, *Y_ = nullptr, *R_ = nullptr;
    std::vector<double*> V_;
    __bf16 *Xh_ = nullptr, *Xl_ = nullptr;
    float* P_ = nullptr;
    int ksplit_ = 1;
    int kb_ct_ = 2;                             // k_gemm_bf16x3_k column tiles per wave
    // stored K row strides (elements); a line of padding measured no faster
    int64_t ldd() const { return Np_; }
    int64_t ldb() const { return Np_; }
    double* Kd_ = nullptr;                      // stored K (null: generated per use)
    __bf16 *Kh_ = nullptr, *Kl_ = nullptr;      // its bf16 parts (split-BF16 handles)
    // packed contractions (stored K only) and the instance selection of the
    // per-instance kernels (NUMA_HIP_CHANDRA_PACK=0 at creation turns both
    // off): instance list, its pinned staging and last upload, the packed
    // operand / result blocks, per-instance combine term counts
    bool lean_ = true;   // packed contractions and instance selection (NUMA_HIP_CHANDRA_PACK)
    bool pack_ = false;
    int64_t nsel_ = -1;  // select_instances(): instances of the per-instance kernels (-1: all)
    int* gmap_ = nullptr;
    int* gmap_h_ = nullptr;
    std::vector<int> gmap_last_;
    double *Xg_ = nullptr, *Yg_ = nullptr;
    float* Pg_ = nullptr;
    int* kcnt_ = nullptr;
    int64_t nsc_ = 0;
    double* sd_ = nullptr;
    double* hs_ = nullptr;
    const double** vptr_ = nullptr;
    double* coef_ = nullptr;
    std::vector<void*> allocs_;
};

// Lockstep GMRES(m) for all active instances on J(X_) z = rhs into sol
// (gmres.hpp:61-223 per instance; IdentityPrecond, whose applications are
// counted as the reference counts them).
void ChandraBatch::gmres_batch(const std::vector<char>& active, const std::vector<double>& rtol,
                               int64_t restart, int64_t max_iters,
                               std::vector<numa_gmres_result>& out, const double* rhs,
                               double* sol) {
    const int64_t B = B_;
    const size_t m = (size_t)restart;
    out.assign((size_t)B, numa_gmres_result{});
    select_instances(nullptr);
    struct AllOnExit {  // every instance again when this returns or throws
        ChandraBatch& h;
        ~AllOnExit() { h.select_instances(nullptr); }
    } all_on_exit{*this};
    // x = 0 for everybody (the Newton caller passes z = 0)
    NUMA_HIP_CHECK(hipMemsetAsync(sol, 0, (size_t)Bp_ * Np_ * sizeof(double), c_.stream));
    std::vector<double> bnorm((size_t)B), beta((size_t)B);
    bdot(rhs, rhs, BEpi{EPI_SQRT, nullptr, MISC(1), nullptr});
    read(MISC(1), B, bnorm.data());
    std::vector<char> done((size_t)B, 1);
    for (int64_t b = 0; b < B; ++b) {
        if (!active[b]) continue;
        if (bnorm[b] == 0.0) {
            out[b].converged = 1;
            continue;
        }
        if (restart == 0) {
            out[b].rel_res = 1.0;
            out[b].breakdown = 1;
            continue;
        }
        if (max_iters == 0) {
            out[b].rel_res = 1.0;
            continue;
        }
        done[b] = 0;
    }
    std::vector<std::vector<std::vector<double>>> Hh((size_t)B);
    std::vector<std::vector<double>> cs((size_t)B, std::vector<double>(m)), sn = cs,
        g((size_t)B, std::vector<double>(m + 1)), sv((size_t)B, std::vector<double>(m + 1));
    for (auto& h : Hh) h.assign(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> col((m + 2) * (size_t)Bp_);  // H columns are read with the padded stride
    bool first = true;
    auto any = [&](const std::vector<char>& f) {
        return std::any_of(f.begin(), f.end(), [](char v) { return v != 0; });
    };
    std::vector<char> notdone((size_t)B);
    for (int64_t b = 0; b < B; ++b) notdone[b] = !done[b];
    while (any(notdone)) {
        // cycle start: r = rhs (first) or rhs - J sol
        if (first) {
            NUMA_HIP_CHECK(hipMemcpyAsync(V_[0], rhs, (size_t)Bp_ * Np_ * sizeof(double),
                                          hipMemcpyDeviceToDevice, c_.stream));
            bdot(V_[0], V_[0], BEpi{EPI_NORM_SCALE, nullptr, MISC(2), S(0)});
        } else {
            const BEpi ep{EPI_NORM_SCALE, nullptr, MISC(2), S(0)};
            select_instances(&notdone);
            jv(sol, nullptr, rhs, V_[0], &ep, &notdone);
            select_instances(nullptr);
            for (int64_t b = 0; b < B; ++b)
                if (notdone[b]) ++out[b].matvecs;
        }
        read(MISC(2), B, beta.data());
        std::vector<char> cyc((size_t)B, 0);
        std::vector<size_t> k((size_t)B, 0);
        for (int64_t b = 0; b < B; ++b) {
            if (!notdone[b]) continue;
            out[b].rel_res = beta[b] / bnorm[b];
            if (out[b].rel_res <= rtol[b]) {
                out[b].converged = 1;
                notdone[b] = 0;
                continue;
            }
            for (auto& row : Hh[b]) std::fill(row.begin(), row.end(), 0.0);
            std::fill(g[b].begin(), g[b].end(), 0.0);
            g[b][0] = beta[b];
            sv[b][0] = beta[b] > 0.0 ? 1.0 / beta[b] : 0.0;
            cyc[b] = out[b].iters < max_iters;
        }
        first = false;
        for (size_t j = 0; j < m && any(cyc); ++j) {
            // w = s_j J V_j into V_{j+1}; MGS
            select_instances(&cyc);
            // w = s_j J V_j into V_{j+1} and the first MGS dot s_0 <w, V_0raw>
            const BEpi h0{EPI_STORE, S(0), H(0, j), nullptr};
            jv(V_[j], S(j), nullptr, V_[j + 1], &h0, &cyc, V_[0]);
            for (size_t i = 0; i <= j; ++i) {
                const bool last = i == j;
                hipLaunchKernelGGL(k_bmgs, sgrid(), dim3(256), 0, c_.stream, (int)N_, (int)Np_,
                                   V_[j + 1], (const double*)V_[i], (const double*)H(i, j),
                                   (const double*)S(i), last ? nullptr : (const double*)V_[i + 1],
                                   c_.red, imap());
                NUMA_HIP_LAUNCH_CHECK();
                bfinish(last ? BEpi{EPI_NORM_SCALE, nullptr, H(j + 1, j), S(j + 1)}
                             : BEpi{EPI_STORE, S(i + 1), H(i + 1, j), nullptr});
            }
            select_instances(nullptr);
            read(H(0, j), (int64_t)(j + 2) * Bp_, col.data());
            for (int64_t b = 0; b < B; ++b) {
                if (!cyc[b]) continue;
                auto& Hb = Hh[b];
                for (size_t i = 0; i <= j + 1; ++i) Hb[i][j] = col[i * Bp_ + b];
                sv[b][j + 1] = Hb[j + 1][j] > 0.0 ? 1.0 / Hb[j + 1][j] : 0.0;
                ++out[b].precs;
                ++out[b].matvecs;
                for (size_t i = 0; i < j; ++i) {
                    const double t = cs[b][i] * Hb[i][j] + sn[b][i] * Hb[i + 1][j];
                    Hb[i + 1][j] = -sn[b][i] * Hb[i][j] + cs[b][i] * Hb[i + 1][j];
                    Hb[i][j] = t;
                }
                const double denom = std::hypot(Hb[j][j], Hb[j + 1][j]);
                if (!(denom > 0.0) || !std::isfinite(denom)) {
                    out[b].breakdown = 1;
                    cyc[b] = 0;
                    notdone[b] = 0;
                    k[b] = 0;  // the reference returns before updating x
                    continue;
                }
                cs[b][j] = Hb[j][j] / denom;
                sn[b][j] = Hb[j + 1][j] / denom;
                Hb[j][j] = denom;
                Hb[j + 1][j] = 0.0;
                g[b][j + 1] = -sn[b][j] * g[b][j];
                g[b][j] = cs[b][j] * g[b][j];
                ++k[b];
                ++out[b].iters;
                out[b].rel_res = std::abs(g[b][j + 1]) / bnorm[b];
                if (out[b].rel_res <= rtol[b] || out[b].iters >= max_iters) cyc[b] = 0;
            }
        }
        // back substitution and x += V (y s) per instance (gmres.hpp:185-202)
        std::vector<double> coef((m + 1) * (size_t)B, 0.0);
        size_t kmax = 0;
        for (int64_t b = 0; b < B; ++b) {
            if (!notdone[b] || out[b].breakdown) continue;
            const size_t kb = k[b];
            std::vector<double> y(kb, 0.0);
            bool bad = false;
            for (size_t ii = kb; ii-- > 0;) {
                const double pivot = Hh[b][ii][ii];
                if (!(std::abs(pivot) > 0.0) || !std::isfinite(pivot)) {
                    bad = true;
                    break;
                }
                double acc = g[b][ii];
                for (size_t jj = ii + 1; jj < kb; ++jj) acc -= Hh[b][ii][jj] * y[jj];
                y[ii] = acc / pivot;
            }
            if (bad) {
                out[b].breakdown = 1;
                notdone[b] = 0;
                continue;
            }
            for (size_t i = 0; i < kb; ++i) coef[i * B + b] = y[i] * sv[b][i];
            kmax = std::max(kmax, kb);
            ++out[b].precs;  // M^{-1}(V y) with the identity
        }
        if (kmax > 0) {
            std::vector<int> kc((size_t)B, 0);
            for (int64_t b = 0; b < B; ++b)
                if (notdone[b] && !out[b].breakdown) kc[b] = (int)k[b];
            NUMA_HIP_CHECK(hipMemcpyAsync(coef_, coef.data(), kmax * B * sizeof(double),
                                          hipMemcpyHostToDevice, c_.stream));
            NUMA_HIP_CHECK(hipMemcpyAsync(kcnt_, kc.data(), B * sizeof(int),
                                          hipMemcpyHostToDevice, c_.stream));
            hipLaunchKernelGGL(k_ch_combine, bgrid(), dim3(256), 0, c_.stream, (int)N_, (int)Np_,
                               (int)B, (const int*)kcnt_, (const double* const*)vptr_,
                               (const double*)coef_, sol);
            NUMA_HIP_LAUNCH_CHECK();
            c_.sync();
        }
        for (int64_t b = 0; b < B; ++b) {
            if (!notdone[b]) continue;
            if (out[b].rel_res <= rtol[b]) {
                out[b].converged = 1;
                notdone[b] = 0;
            } else if (out[b].iters >= max_iters) {
                notdone[b] = 0;
            }
        }
    }
}

// newton_krylov (krylov_newton.hpp:69-114) or chebyshev_inexact (:116-189)
// for every instance, in lockstep.
void ChandraBatch::solve(const numa_krylov_params& p, double* xuser, numa_krylov_result* res) {
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t B = B_;
    require(p.gmres_restart >= 0 && p.gmres_restart <= m_,
            "gmres_restart exceeds the handle's basis capacity");
    for (int64_t b = 0; b < B; ++b) res[b] = numa_krylov_result{};
    pack(xuser, X_, true);
    std::vector<double> rnorm((size_t)B), rprev((size_t)B), eta_prev((size_t)B, p.eta_init),
        eta((size_t)B), lam((size_t)B), fn((size_t)B), acc((size_t)B);
    eval_F(X_, F_, Ax_, MISC(0));
    read(MISC(0), B, rnorm.data());
    std::vector<char> active((size_t)B, 1);
    for (int64_t b = 0; b < B; ++b) {
        res[b].n_F = 1;
        res[b].res0 = rnorm[b];
        rprev[b] = rnorm[b];
    }
    std::vector<numa_gmres_result> g;
    for (int64_t k = 0; k < p.max_nl; ++k) {
        bool any = false;
        for (int64_t b = 0; b < B; ++b) {
            if (!active[b]) continue;
            if (nl_converged(p, rnorm[b], res[b].res0, res[b].iters, rprev[b])) {
                res[b].converged = 1;
                active[b] = 0;
                continue;
            }
            any = true;
            // krylov_newton.hpp:54-65
            double e;
            if (p.forcing == NUMA_FORCING_CONSTANT || k == 0) {
                e = p.eta_init;
            } else {
                e = p.gamma * std::pow(rnorm[b] / rprev[b], p.alpha);
                const double safe = p.gamma * std::pow(eta_prev[b], p.alpha);
                if (safe > 0.1) e = std::max(e, safe);
                e = std::min(e, p.eta_max);
                e = std::max(e, 0.5 * p.rtol_nl * res[b].res0 / rnorm[b]);
                e = std::min(e, p.eta_max);
            }
            eta[b] = tight_cap(p, e, rnorm[b], res[b].res0);
        }
        if (!any) break;
        gmres_batch(active, eta, p.gmres_restart, p.gmres_max_iters, g, F_, Z_);
        std::vector<char> pending((size_t)B, 0);
        for (int64_t b = 0; b < B; ++b) {
            if (!active[b]) continue;
            res[b].n_JVP += g[b].matvecs;
            res[b].n_lin_iters += g[b].iters;
            res[b].n_prec += g[b].precs;
            if (!g[b].converged) ++res[b].n_lin_failures;
            if (g[b].breakdown) {
                active[b] = 0;
                continue;
            }
            lam[b] = 1.0;
            pending[b] = 1;
        }
        std::vector<char> accepted((size_t)B, 0);
        auto any_of = [](const std::vector<char>& f) {
            return std::any_of(f.begin(), f.end(), [](char v) { return v != 0; });
        };
        auto select_trial = [&](const std::vector<double>& take) {
            write(MISC(4), take);
            for (auto pr : {std::pair<double*, double*>{Xt_, X_}, {Ft_, F_}, {Axt_, Ax_}}) {
                hipLaunchKernelGGL(k_ch_select, bgrid(), dim3(256), 0, c_.stream, (int)N_,
                                   (int)Np_, (const double*)MISC(4), (const double*)pr.first,
                                   pr.second);
                NUMA_HIP_LAUNCH_CHECK();
            }
        };
        if (method_ == NUMA_METHOD_CHEBYSHEV && any_of(pending)) {
            // chebyshev_inexact (krylov_newton.hpp:147-169): J u = F''[z, z] with
            // eta2 = min(eta2_max, eta2_factor eta1), then the full step
            // x - z - u/2 if it lowers ||F||, else the Newton step with backtracking
            d2(Z_, W_, &pending);
            std::vector<double> eta2((size_t)B, 0.0);
            for (int64_t b = 0; b < B; ++b) {
                if (!pending[b]) continue;
                ++res[b].n_D2;
                eta2[b] = std::min(p.eta2_max, p.eta2_factor * eta[b]);
            }
            std::vector<numa_gmres_result> g2;
            gmres_batch(pending, eta2, p.gmres_restart, p.gmres_max_iters, g2, W_, U_);
            std::vector<char> full((size_t)B, 0);
            for (int64_t b = 0; b < B; ++b) {
                if (!pending[b]) continue;
                res[b].n_JVP += g2[b].matvecs;
                res[b].n_lin_iters += g2[b].iters;
                res[b].n_prec += g2[b].precs;
                if (!g2[b].converged) ++res[b].n_lin_failures;
                full[b] = !g2[b].breakdown;
            }
            if (any_of(full)) {
                hipLaunchKernelGGL(k_ch_cheb_trial, bgrid(), dim3(256), 0, c_.stream, (int)N_,
                                   (int)Np_, (const double*)X_, (const double*)Z_,
                                   (const double*)U_, Xt_);
                NUMA_HIP_LAUNCH_CHECK();
                eval_F(Xt_, Ft_, Axt_, MISC(0), &full);
                read(MISC(0), B, fn.data());
                std::vector<double> take((size_t)Bp_, 0.0);
                for (int64_t b = 0; b < B; ++b) {
                    if (!full[b]) continue;
                    ++res[b].n_F;
                    if (fn[b] < rnorm[b]) {
                        pending[b] = 0;
                        accepted[b] = 1;
                        take[b] = 1.0;
                        acc[b] = fn[b];
                    }
                }
                select_trial(take);
            }
        }
        // Armijo backtracking (krylov_newton.hpp:96-104), all pending instances at once
        while (std::any_of(pending.begin(), pending.end(), [](char v) { return v != 0; })) {
            std::vector<double> l((size_t)Bp_, 0.0);
            for (int64_t b = 0; b < B; ++b) l[b] = pending[b] ? lam[b] : 0.0;
            write(MISC(3), l);
            hipLaunchKernelGGL(k_ch_trial, bgrid(), dim3(256), 0, c_.stream, (int)N_, (int)Np_,
                               (const double*)X_, (const double*)Z_, (const double*)MISC(3), Xt_);
            NUMA_HIP_LAUNCH_CHECK();
            eval_F(Xt_, Ft_, Axt_, MISC(0), &pending);
            read(MISC(0), B, fn.data());
            std::vector<double> take((size_t)Bp_, 0.0);
            for (int64_t b = 0; b < B; ++b) {
                if (!pending[b]) continue;
                ++res[b].n_F;
                if (fn[b] <= (1.0 - p.armijo * lam[b]) * rnorm[b]) {
                    pending[b] = 0;
                    accepted[b] = 1;
                    take[b] = 1.0;
                    acc[b] = fn[b];
                } else {
                    lam[b] *= 0.5;
                    if (lam[b] < p.min_lambda) {
                        pending[b] = 0;
                        active[b] = 0;  // line search failed: stop this instance
                    }
                }
            }
            select_trial(take);
        }
        for (int64_t b = 0; b < B; ++b) {
            if (!accepted[b]) continue;
            rprev[b] = rnorm[b];
            rnorm[b] = acc[b];
            eta_prev[b] = eta[b];
            ++res[b].iters;
        }
    }
    pack(X_, xuser, false);
    c_.sync();
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int64_t b = 0; b < B; ++b) {
        if (rnorm[b] <= p.rtol_nl * res[b].res0) res[b].converged = 1;
        res[b].res = rnorm[b];
        res[b].t_total = t;
    }
}

}  // namespace numa_hip

namespace {
// ---- measured dense MFMA throughput ----------------------------------------
// (in this file: built with the accumulators in VGPRs, as the contractions)
// Each wave runs kPeakChains independent accumulator chains of one MFMA kind
// for `iters` steps (no memory traffic; the sums go to `sink` only in a case
// that does not occur, which keeps every product live).
typedef __bf16 peak_bf16x8 __attribute__((ext_vector_type(8)));
typedef float peak_floatx16 __attribute__((ext_vector_type(16)));
typedef double peak_doublex4 __attribute__((ext_vector_type(4)));
constexpr int kPeakThreads = 256;

template <int KIND, int kPeakChains>
__global__ __launch_bounds__(kPeakThreads) void k_mfma_peak(int iters, double* sink) {
    const int t = (int)threadIdx.x;
    double s = 0.0;
    if constexpr (KIND == 0) {
        peak_bf16x8 a, b;
        for (int e = 0; e < 8; ++e) {
            a[e] = static_cast<__bf16>(1e-3f * (float)(t + e));
            b[e] = static_cast<__bf16>(2e-3f * (float)(t - e));
        }
        peak_floatx16 acc[kPeakChains] = {};
#pragma unroll 2
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int q = 0; q < kPeakChains; ++q)
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
        for (int q = 0; q < kPeakChains; ++q)
            for (int r = 0; r < 16; ++r) s += acc[q][r];
    } else {
        const double a = 1e-3 * t, b = 2e-3 * (t + 1);
        peak_doublex4 acc[kPeakChains] = {};
#pragma unroll 2
        for (int i = 0; i < iters; ++i)
#pragma unroll
            for (int q = 0; q < kPeakChains; ++q)
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
        for (int q = 0; q < kPeakChains; ++q)
            for (int r = 0; r < 4; ++r) s += acc[q][r];
    }
    if (s == -1234.5) sink[blockIdx.x] = s;
}

}  // namespace

struct numa_hip_chandra {
    numa_hip::Ctx* ctx;
    std::unique_ptr<numa_hip::ChandraBatch> impl;
};

using numa_hip::Error;


extern "C" {

int numa_hip_mfma_peak(numa_hip_ctx* ctx, int64_t kind, double* tflops) {
    NUMA_API_BEGIN
    numa_hip::require(ctx && ctx->impl, "null context");
    NUMA_HIP_CHECK(hipSetDevice(ctx->impl->device));
    numa_hip::Ctx& c = *ctx->impl;
    numa_hip::require(kind == 0 || kind == 1, "kind must be 0 (bf16) or 1 (fp64)");
    numa_hip::require(tflops, "null result pointer");
    const double flop_per_mfma = kind == 0 ? 2.0 * 32 * 32 * 16 : 2.0 * 16 * 16 * 4;
    hipEvent_t e0, e1;
    NUMA_HIP_CHECK(hipEventCreate(&e0));
    NUMA_HIP_CHECK(hipEventCreate(&e1));
    double best_tf = 0.0;
    // 4 or 8 accumulator chains per wave, 4 or 8 workgroups of 4 waves per CU
    for (int chains : {4, 8})
        for (int per_cu : {4, 8}) {
            const unsigned blocks = (unsigned)c.cu_count * per_cu;
            const int iters = (kind == 0 ? 8192 : 4096) / chains * 4 / per_cu * 2;
            const double flop =
                flop_per_mfma * chains * (double)iters * (kPeakThreads / 64) * (double)blocks;
            float best = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {  // the first is a warm-up
                NUMA_HIP_CHECK(hipEventRecord(e0, c.stream));
                auto go = [&](auto kern) {
                    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kPeakThreads), 0, c.stream, iters,
                                       c.red.partials);
                };
                if (kind == 0)
                    chains == 4 ? go(k_mfma_peak<0, 4>) : go(k_mfma_peak<0, 8>);
                else
                    chains == 4 ? go(k_mfma_peak<1, 4>) : go(k_mfma_peak<1, 8>);
                NUMA_HIP_LAUNCH_CHECK();
                NUMA_HIP_CHECK(hipEventRecord(e1, c.stream));
                NUMA_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                NUMA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0) best = std::min(best, ms);
            }
            best_tf = std::max(best_tf, flop / (best * 1e-3) / 1e12);
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *tflops = best_tf;
    NUMA_API_END
}


int numa_hip_chandra_create(numa_hip_ctx* ctx, int64_t n, int64_t batch, const double* c_host,
                            int64_t rule, int64_t form, int64_t precision, int64_t gmres_restart,
                            numa_hip_chandra** out) {
    NUMA_API_BEGIN
    numa_hip::require(ctx && ctx->impl && c_host && out, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(ctx->impl->device));
    auto h = std::make_unique<numa_hip_chandra>();
    h->ctx = ctx->impl;
    h->impl = std::make_unique<numa_hip::ChandraBatch>(*ctx->impl, n, batch, c_host, (int)rule,
                                                       (int)form, (int)precision, gmres_restart);
    ++ctx->impl->live_handles;
    *out = h.release();
    NUMA_API_END
}

int numa_hip_chandra_destroy(numa_hip_chandra* h) {
    NUMA_API_BEGIN
    if (h) {
        NUMA_HIP_CHECK(hipStreamSynchronize(h->ctx->stream));
        --h->ctx->live_handles;
        delete h;
    }
    NUMA_API_END
}

int numa_hip_chandra_eval(numa_hip_chandra* h, const double* x_dev, double* F_dev,
                          double* norms_host) {
    NUMA_API_BEGIN
    numa_hip::require(h && x_dev && F_dev, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    h->impl->eval_user(x_dev, F_dev, norms_host);
    NUMA_API_END
}

int numa_hip_chandra_jvp(numa_hip_chandra* h, const double* x_dev, const double* v_dev,
                         double* out_dev) {
    NUMA_API_BEGIN
    numa_hip::require(h && x_dev && v_dev && out_dev, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    h->impl->jvp_user(x_dev, v_dev, out_dev);
    NUMA_API_END
}

int numa_hip_chandra_d2(numa_hip_chandra* h, const double* x_dev, const double* v_dev,
                        double* out_dev) {
    NUMA_API_BEGIN
    numa_hip::require(h && x_dev && v_dev && out_dev, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    h->impl->d2_user(x_dev, v_dev, out_dev);
    NUMA_API_END
}

int numa_hip_chandra_set_method(numa_hip_chandra* h, int64_t method) {
    NUMA_API_BEGIN
    numa_hip::require(h != nullptr, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    h->impl->set_method((int)method);
    NUMA_API_END
}

int numa_hip_chandra_quadrature(int64_t n, int64_t rule, double* nodes_host,
                                double* weights_host) {
    NUMA_API_BEGIN
    numa_hip::require(n > 0 && nodes_host && weights_host, "bad argument");
    numa_hip::require(rule == 0 || rule == 1, "rule must be 0 (Gauss-Legendre) or 1 (midpoint)");
    std::vector<double> mu, w;
    if (rule == 0) {
        numa_hip::gauss_legendre_unit((size_t)n, mu, w);
    } else {
        mu.resize((size_t)n);
        w.assign((size_t)n, 1.0 / double(n));
        for (int64_t i = 0; i < n; ++i) mu[(size_t)i] = (double(i) + 0.5) / double(n);
    }
    std::copy(mu.begin(), mu.end(), nodes_host);
    std::copy(w.begin(), w.end(), weights_host);
    NUMA_API_END
}

int numa_hip_chandra_solve(numa_hip_chandra* h, const numa_krylov_params* p, double* x_dev,
                           numa_krylov_result* results) {
    NUMA_API_BEGIN
    numa_hip::require(h && p && x_dev && results, "null argument");
    numa_hip::require(p->max_nl >= 0 && p->gmres_max_iters >= 0, "negative iteration limits");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    h->impl->solve(*p, x_dev, results);
    NUMA_API_END
}

int numa_hip_chandra_gemm_cost(numa_hip_chandra* h, int64_t precision, int64_t reps,
                               double* ms_per_call, double* flops_per_call) {
    NUMA_API_BEGIN
    numa_hip::require(h && ms_per_call, "null argument");
    NUMA_HIP_CHECK(hipSetDevice(h->ctx->device));
    *ms_per_call = h->impl->gemm_cost((int)precision, reps, flops_per_call);
    NUMA_API_END
}

}  // extern "C"
