// blas.hpp - BLAS-1 and Arnoldi kernels (operators/operator.hpp:19-41,
// linalg/gmres.hpp:141-153) on gfx950.
//
// Grid-stride streaming kernels: per iteration a thread handles two element
// pairs, q and q + 256, so each wave-wide access is one contiguous span
// (16-byte FP64 / 8-byte FP32 per lane) and two of them are in flight per
// stream.  All loads of an iteration are issued before its stores.  The launch
// shape is a fixed function of the length, so reductions sum the same
// partials in the same order (bitwise repeatable); dots accumulate in FP64
// whatever the storage type.  Pointers that are not pair-aligned take the
// scalar-load variant of the same kernel.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace numa_hip {

template <class T>
__device__ inline double dv(const T* p, int64_t i) {
    return static_cast<double>(p[i]);
}

template <class T>
struct alignas(2 * sizeof(T)) Pair {
    T a, b;
};

template <bool VEC, class T>
__device__ inline Pair<T> ldp(const T* p, int64_t q) {
    if constexpr (VEC) {
        return reinterpret_cast<const Pair<T>*>(p)[q];
    } else {
        return Pair<T>{p[2 * q], p[2 * q + 1]};
    }
}

template <bool VEC, class T>
__device__ inline void stp(T* p, int64_t q, const Pair<T>& v) {
    if constexpr (VEC) {
        reinterpret_cast<Pair<T>*>(p)[q] = v;
    } else {
        p[2 * q] = v.a;
        p[2 * q + 1] = v.b;
    }
}

// Visits the pairs of [0, N): f2(q0, q1, ok0, ok1) with q1 = q0 + 256, then
// the odd last element (if any) with f1(N-1) on one thread.  g is the block's
// position in its (possibly batched) launch.
template <class F2, class F1>
__device__ inline void for_each_pair(GridPos g, int64_t N, F2&& f2, F1&& f1) {
    const int64_t P = N >> 1;
    const int64_t step = (int64_t)g.nb * (2 * kBlock);
    for (int64_t base = (int64_t)g.bid * (2 * kBlock); base < P; base += step) {
        const int64_t q0 = base + threadIdx.x, q1 = q0 + kBlock;
        f2(q0, q1, q0 < P, q1 < P);
    }
    if ((N & 1) && g.bid == 0 && threadIdx.x == 0) f1(N - 1);
}
template <class F2, class F1>
__device__ inline void for_each_pair(int64_t N, F2&& f2, F1&& f1) {
    for_each_pair(hw_grid(), N, f2, f1);
}

inline bool aligned2(const void* p, size_t elem) {
    return p == nullptr || (reinterpret_cast<uintptr_t>(p) % (2 * elem)) == 0;
}

// sum_i a_i b_i (b may equal a) -> one partial per block.
template <bool VEC, class TA, class TB>
__device__ inline void dot_body(GridPos g, int64_t N, const TA* a, const TB* b, Reduce red) {
    double acc = 0.0;
    for_each_pair(
        g, N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            Pair<TA> x0{}, x1{};
            Pair<TB> y0{}, y1{};
            if (ok0) {
                x0 = ldp<VEC>(a, q0);
                y0 = ldp<VEC>(b, q0);
            }
            if (ok1) {
                x1 = ldp<VEC>(a, q1);
                y1 = ldp<VEC>(b, q1);
            }
            acc += (double)x0.a * (double)y0.a;
            acc += (double)x0.b * (double)y0.b;
            acc += (double)x1.a * (double)y1.a;
            acc += (double)x1.b * (double)y1.b;
        },
        [&](int64_t i) { acc += dv(a, i) * dv(b, i); });
    block_partial<kBlock>(acc, red, g.bid);
}
template <bool VEC, class TA, class TB>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t N, const TA* a, const TB* b, Reduce red) {
    dot_body<VEC, TA, TB>(hw_grid(), N, a, b, red);
}

template <class TA, class TB>
void launch_dot(Ctx& c, int64_t N, const TA* a, const TB* b, const Epilogue& ep) {
    const unsigned nb = stream_blocks(N);
    if (aligned2(a, sizeof(TA)) && aligned2(b, sizeof(TB)))
        hipLaunchKernelGGL((k_dot<true, TA, TB>), dim3(nb), dim3(kBlock), 0, c.stream, N, a, b,
                           c.red);
    else
        hipLaunchKernelGGL((k_dot<false, TA, TB>), dim3(nb), dim3(kBlock), 0, c.stream, N, a, b,
                           c.red);
    NUMA_HIP_LAUNCH_CHECK();
    c.finish(nb, ep);
}

// Fault injection: element 0 of v and the scalar slots s1, s2 (each may be
// null) become NaN.  One thread.
template <class T>
__global__ void k_poison(T* v, double* s1, double* s2) {
    const double nan = __builtin_nan("");
    if (v) v[0] = static_cast<T>(nan);
    if (s1) s1[0] = nan;
    if (s2) s2[0] = nan;
}

template <class T>
void launch_poison(Ctx& c, T* v, double* s1, double* s2) {
    hipLaunchKernelGGL((k_poison<T>), dim3(1), dim3(1), 0, c.stream, v, s1, s2);
    NUMA_HIP_LAUNCH_CHECK();
}

// y += a x  (operator.hpp:32-41)
template <bool VEC, class TX, class TY>
__global__ __launch_bounds__(kBlock) void k_axpy(int64_t N, double a, const TX* x, TY* y) {
#pragma clang fp contract(off)
    auto upd = [a](Pair<TY> yv, Pair<TX> xv) {
        return Pair<TY>{static_cast<TY>((double)yv.a + a * (double)xv.a),
                        static_cast<TY>((double)yv.b + a * (double)xv.b)};
    };
    for_each_pair(
        N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            Pair<TX> x0{}, x1{};
            Pair<TY> y0{}, y1{};
            if (ok0) {
                x0 = ldp<VEC>(x, q0);
                y0 = ldp<VEC>(y, q0);
            }
            if (ok1) {
                x1 = ldp<VEC>(x, q1);
                y1 = ldp<VEC>(y, q1);
            }
            if (ok0) stp<VEC>(y, q0, upd(y0, x0));
            if (ok1) stp<VEC>(y, q1, upd(y1, x1));
        },
        [&](int64_t i) { y[i] = static_cast<TY>(dv(y, i) + a * dv(x, i)); });
}

template <bool VEC, class T>
__device__ inline void fill_body(GridPos g, int64_t N, T v, T* y) {
    for_each_pair(
        g, N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            if (ok0) stp<VEC>(y, q0, Pair<T>{v, v});
            if (ok1) stp<VEC>(y, q1, Pair<T>{v, v});
        },
        [&](int64_t i) { y[i] = v; });
}
template <bool VEC, class T>
__global__ __launch_bounds__(kBlock) void k_fill(int64_t N, T v, T* y) {
    fill_body<VEC, T>(hw_grid(), N, v, y);
}

template <bool VEC, class TI, class TO>
__device__ inline void convert_body(GridPos g, int64_t N, const TI* x, TO* y) {
    for_each_pair(
        g, N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            Pair<TI> a0{}, a1{};
            if (ok0) a0 = ldp<VEC>(x, q0);
            if (ok1) a1 = ldp<VEC>(x, q1);
            if (ok0) stp<VEC>(y, q0, Pair<TO>{static_cast<TO>(a0.a), static_cast<TO>(a0.b)});
            if (ok1) stp<VEC>(y, q1, Pair<TO>{static_cast<TO>(a1.a), static_cast<TO>(a1.b)});
        },
        [&](int64_t i) { y[i] = static_cast<TO>(x[i]); });
}
template <bool VEC, class TI, class TO>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t N, const TI* x, TO* y) {
    convert_body<VEC, TI, TO>(hw_grid(), N, x, y);
}

// One fused modified-Gram-Schmidt pass (gmres.hpp:141-144):
//   w -= h_i * (s_i V_i)            with h_i = *hi (already the true dot)
//   !LAST: partial of sum w * V_{i+1}   (-> H[i+1][j] = s_{i+1} * sum)
//   LAST:  partial of sum w * w         (-> H[j+1][j] = sqrt, s_{j+1} = 1/sqrt)
// V_i are stored unnormalised with scales s_i.
template <bool VEC, bool LAST, class T>
__device__ inline void mgs_body(GridPos g, int64_t N, T* w, const T* Vi, const double* hi,
                                const double* si, const T* Vnext, Reduce red) {
#pragma clang fp contract(off)
    const double coef = (*hi) * (*si);
    double acc = 0.0;
    auto upd = [coef](Pair<T> wv, Pair<T> vv) {
        return Pair<T>{static_cast<T>((double)wv.a - coef * (double)vv.a),
                       static_cast<T>((double)wv.b - coef * (double)vv.b)};
    };
    auto dot = [](Pair<T> u, Pair<T> v) {
        return (double)u.a * (double)v.a + (double)u.b * (double)v.b;
    };
    for_each_pair(
        g, N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            Pair<T> w0{}, w1{}, v0{}, v1{}, n0{}, n1{};
            if (ok0) {
                w0 = ldp<VEC>(w, q0);
                v0 = ldp<VEC>(Vi, q0);
                if constexpr (!LAST) n0 = ldp<VEC>(Vnext, q0);
            }
            if (ok1) {
                w1 = ldp<VEC>(w, q1);
                v1 = ldp<VEC>(Vi, q1);
                if constexpr (!LAST) n1 = ldp<VEC>(Vnext, q1);
            }
            w0 = upd(w0, v0);
            w1 = upd(w1, v1);
            if (ok0) stp<VEC>(w, q0, w0);
            if (ok1) stp<VEC>(w, q1, w1);
            if constexpr (LAST) {
                acc += dot(w0, w0);
                acc += dot(w1, w1);
            } else {
                acc += dot(w0, n0);
                acc += dot(w1, n1);
            }
        },
        [&](int64_t i) {
            const T wn = static_cast<T>(dv(w, i) - coef * dv(Vi, i));
            w[i] = wn;
            const double wd = static_cast<double>(wn);
            acc += LAST ? wd * wd : wd * dv(Vnext, i);
        });
    block_partial<kBlock>(acc, red, g.bid);
}
template <bool VEC, bool LAST, class T>
__global__ __launch_bounds__(kBlock) void k_mgs(int64_t N, T* w, const T* Vi, const double* hi,
                                                 const double* si, const T* Vnext, Reduce red) {
    mgs_body<VEC, LAST, T>(hw_grid(), N, w, Vi, hi, si, Vnext, red);
}

// y (=|+=) sum_k c_k X_k over up to kMaxComb vectors (x += V y, gmres.hpp:198-202).
constexpr int kMaxComb = 32;
template <class T>
struct CombArgs {
    const T* v[kMaxComb];
    double c[kMaxComb];
    int count;
};

template <bool VEC, class T, class TY>
__device__ inline void combine_body(GridPos g, int64_t N, const CombArgs<T>& a, TY* y,
                                    int accumulate) {
#pragma clang fp contract(off)
    for_each_pair(
        g, N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            double s0a = 0.0, s0b = 0.0, s1a = 0.0, s1b = 0.0;
            if (accumulate) {
                if (ok0) {
                    const Pair<TY> t = ldp<VEC>(y, q0);
                    s0a = (double)t.a;
                    s0b = (double)t.b;
                }
                if (ok1) {
                    const Pair<TY> t = ldp<VEC>(y, q1);
                    s1a = (double)t.a;
                    s1b = (double)t.b;
                }
            }
            for (int k = 0; k < a.count; ++k) {
                Pair<T> x0{}, x1{};
                if (ok0) x0 = ldp<VEC>(a.v[k], q0);
                if (ok1) x1 = ldp<VEC>(a.v[k], q1);
                s0a += a.c[k] * (double)x0.a;
                s0b += a.c[k] * (double)x0.b;
                s1a += a.c[k] * (double)x1.a;
                s1b += a.c[k] * (double)x1.b;
            }
            if (ok0) stp<VEC>(y, q0, Pair<TY>{static_cast<TY>(s0a), static_cast<TY>(s0b)});
            if (ok1) stp<VEC>(y, q1, Pair<TY>{static_cast<TY>(s1a), static_cast<TY>(s1b)});
        },
        [&](int64_t i) {
            double s = accumulate ? dv(y, i) : 0.0;
            for (int k = 0; k < a.count; ++k) s += a.c[k] * dv(a.v[k], i);
            y[i] = static_cast<TY>(s);
        });
}
template <bool VEC, class T, class TY>
__global__ __launch_bounds__(kBlock) void k_combine(int64_t N, CombArgs<T> a, TY* y,
                                                     int accumulate) {
    combine_body<VEC, T, TY>(hw_grid(), N, a, y, accumulate);
}

// z = r / d   (BratuJacobi::Bound, bratu3d.hpp:245-248)
template <bool VEC, class T>
__global__ __launch_bounds__(kBlock) void k_diag_solve(int64_t N, const T* r, const double* d,
                                                        T* z) {
    for_each_pair(
        N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            Pair<T> r0{}, r1{};
            Pair<double> d0{1.0, 1.0}, d1{1.0, 1.0};
            if (ok0) {
                r0 = ldp<VEC>(r, q0);
                d0 = ldp<VEC>(d, q0);
            }
            if (ok1) {
                r1 = ldp<VEC>(r, q1);
                d1 = ldp<VEC>(d, q1);
            }
            if (ok0)
                stp<VEC>(z, q0, Pair<T>{static_cast<T>((double)r0.a / d0.a),
                                        static_cast<T>((double)r0.b / d0.b)});
            if (ok1)
                stp<VEC>(z, q1, Pair<T>{static_cast<T>((double)r1.a / d1.a),
                                        static_cast<T>((double)r1.b / d1.b)});
        },
        [&](int64_t i) { z[i] = static_cast<T>(dv(r, i) / d[i]); });
}

namespace {
// d = 6 h2inv - lambda e^u   (bratu3d.hpp:59-69)
__global__ __launch_bounds__(kBlock) void k_jacobi_diag(int64_t N, double h2inv, double lambda_v,
                                                        const double* lam_ptr, const double* u,
                                                        double* d) {
#pragma clang fp contract(off)
    const double lambda = lam_ptr ? *lam_ptr : lambda_v;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        d[i] = 6.0 * h2inv - lambda * exp(u[i]);
}

// out = x + a d  (y = x; axpy(a, d, y), operator.hpp:32-41)
__global__ __launch_bounds__(kBlock) void k_axpy_to(int64_t N, const double* x, double a,
                                                    const double* d, double* out) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        out[i] = x[i] + a * d[i];
}

// out = s x
__global__ __launch_bounds__(kBlock) void k_scale_to(int64_t N, double s, const double* x,
                                                     double* out) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        out[i] = s * x[i];
}

// out = x / s
__global__ __launch_bounds__(kBlock) void k_div_to(int64_t N, const double* x, double s,
                                                   double* out) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        out[i] = x[i] / s;
}

// q_i = 1 + 0.1 sin(0.6180339887498948 (i + 1)): the inverse iteration's start
// (analysis/spectral.hpp:51-55)
__global__ __launch_bounds__(kBlock) void k_eig_start(int64_t N, double* q) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        q[i] = 1.0 + 0.1 * sin(0.6180339887498948 * static_cast<double>(i + 1));
}

// out = s ((a - b) / h)  (difference quotient, then scaled; in place allowed)
__global__ __launch_bounds__(kBlock) void k_diffq(int64_t N, const double* a, const double* b,
                                                  double h, double s, double* out) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock) {
        const double q = (a[i] - b[i]) / h;
        out[i] = s * q;
    }
}

// out = -lambda e^u v v   (bratu3d.hpp:46-56)
__global__ __launch_bounds__(kBlock) void k_d2_analytic(int64_t N, double lambda_v,
                                                        const double* lam_ptr, const double* u,
                                                        const double* v, double* out) {
#pragma clang fp contract(off)
    const double lambda = lam_ptr ? *lam_ptr : lambda_v;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        out[i] = -lambda * exp(u[i]) * v[i] * v[i];
}
}  // namespace

template <class T>
void launch_fill(Ctx& c, int64_t N, T v, T* y) {
    if (aligned2(y, sizeof(T)))
        hipLaunchKernelGGL((k_fill<true, T>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N,
                           v, y);
    else
        hipLaunchKernelGGL((k_fill<false, T>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream,
                           N, v, y);
    NUMA_HIP_LAUNCH_CHECK();
}

template <class TI, class TO>
void launch_convert(Ctx& c, int64_t N, const TI* x, TO* y) {
    if (aligned2(x, sizeof(TI)) && aligned2(y, sizeof(TO)))
        hipLaunchKernelGGL((k_convert<true, TI, TO>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, x, y);
    else
        hipLaunchKernelGGL((k_convert<false, TI, TO>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, x, y);
    NUMA_HIP_LAUNCH_CHECK();
}

// y = x (converted) and the partials of sum x_i^2 in one pass; the partials
// are summed in the order of k_dot(x, x), so the epilogue sees the same total.
template <bool VEC, class TI, class TO>
__device__ inline void convert_norm_body(GridPos g, int64_t N, const TI* x, TO* y, Reduce red) {
    double acc = 0.0;
    for_each_pair(
        g, N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            Pair<TI> a0{}, a1{};
            if (ok0) a0 = ldp<VEC>(x, q0);
            if (ok1) a1 = ldp<VEC>(x, q1);
            acc += (double)a0.a * (double)a0.a;
            acc += (double)a0.b * (double)a0.b;
            acc += (double)a1.a * (double)a1.a;
            acc += (double)a1.b * (double)a1.b;
            if (ok0) stp<VEC>(y, q0, Pair<TO>{static_cast<TO>(a0.a), static_cast<TO>(a0.b)});
            if (ok1) stp<VEC>(y, q1, Pair<TO>{static_cast<TO>(a1.a), static_cast<TO>(a1.b)});
        },
        [&](int64_t i) {
            acc += dv(x, i) * dv(x, i);
            y[i] = static_cast<TO>(x[i]);
        });
    block_partial<kBlock>(acc, red, g.bid);
}
template <bool VEC, class TI, class TO>
__global__ __launch_bounds__(kBlock) void k_convert_norm(int64_t N, const TI* x, TO* y,
                                                         Reduce red) {
    convert_norm_body<VEC, TI, TO>(hw_grid(), N, x, y, red);
}

template <class TI, class TO>
void launch_convert_norm(Ctx& c, int64_t N, const TI* x, TO* y, const Epilogue& ep) {
    const unsigned nb = stream_blocks(N);
    if (aligned2(x, sizeof(TI)) && aligned2(y, sizeof(TO)))
        hipLaunchKernelGGL((k_convert_norm<true, TI, TO>), dim3(nb), dim3(kBlock), 0, c.stream, N,
                           x, y, c.red);
    else
        hipLaunchKernelGGL((k_convert_norm<false, TI, TO>), dim3(nb), dim3(kBlock), 0, c.stream,
                           N, x, y, c.red);
    NUMA_HIP_LAUNCH_CHECK();
    c.finish(nb, ep);
}

inline void launch_axpy_to(Ctx& c, int64_t N, const double* x, double a, const double* d,
                           double* out) {
    hipLaunchKernelGGL(k_axpy_to, dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N, x, a, d,
                       out);
    NUMA_HIP_LAUNCH_CHECK();
}

inline void launch_scale_to(Ctx& c, int64_t N, double s, const double* x, double* out) {
    hipLaunchKernelGGL(k_scale_to, dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N, s, x,
                       out);
    NUMA_HIP_LAUNCH_CHECK();
}

inline void launch_div_to(Ctx& c, int64_t N, const double* x, double s, double* out) {
    hipLaunchKernelGGL(k_div_to, dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N, x, s, out);
    NUMA_HIP_LAUNCH_CHECK();
}

inline void launch_eig_start(Ctx& c, int64_t N, double* q) {
    hipLaunchKernelGGL(k_eig_start, dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N, q);
    NUMA_HIP_LAUNCH_CHECK();
}

inline void launch_diffq(Ctx& c, int64_t N, const double* a, const double* b, double h, double s,
                         double* out) {
    hipLaunchKernelGGL(k_diffq, dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N, a, b, h, s,
                       out);
    NUMA_HIP_LAUNCH_CHECK();
}

template <class TX, class TY>
void launch_axpy(Ctx& c, int64_t N, double a, const TX* x, TY* y) {
    if (aligned2(x, sizeof(TX)) && aligned2(y, sizeof(TY)))
        hipLaunchKernelGGL((k_axpy<true, TX, TY>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, a, x, y);
    else
        hipLaunchKernelGGL((k_axpy<false, TX, TY>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, a, x, y);
    NUMA_HIP_LAUNCH_CHECK();
}

template <bool LAST, class T>
void launch_mgs_impl(Ctx& c, int64_t N, T* w, const T* Vi, const double* hi, const double* si,
                     const T* Vnext, unsigned nb) {
    if (aligned2(w, sizeof(T)) && aligned2(Vi, sizeof(T)) && aligned2(Vnext, sizeof(T)))
        hipLaunchKernelGGL((k_mgs<true, LAST, T>), dim3(nb), dim3(kBlock), 0, c.stream, N, w, Vi,
                           hi, si, Vnext, c.red);
    else
        hipLaunchKernelGGL((k_mgs<false, LAST, T>), dim3(nb), dim3(kBlock), 0, c.stream, N, w, Vi,
                           hi, si, Vnext, c.red);
    NUMA_HIP_LAUNCH_CHECK();
}

// MGS pass i; Vnext == null marks the last pass (norm of w).
template <class T>
void launch_mgs(Ctx& c, int64_t N, T* w, const T* Vi, const double* hi, const double* si,
                const T* Vnext, const Epilogue& ep) {
    const unsigned nb = stream_blocks(N);
    if (Vnext)
        launch_mgs_impl<false, T>(c, N, w, Vi, hi, si, Vnext, nb);
    else
        launch_mgs_impl<true, T>(c, N, w, Vi, hi, si, Vnext, nb);
    c.finish(nb, ep);
}

template <class T>
void launch_diag_solve(Ctx& c, int64_t N, const T* r, const double* d, T* z) {
    if (aligned2(r, sizeof(T)) && aligned2(d, sizeof(double)) && aligned2(z, sizeof(T)))
        hipLaunchKernelGGL((k_diag_solve<true, T>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, r, d, z);
    else
        hipLaunchKernelGGL((k_diag_solve<false, T>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, r, d, z);
    NUMA_HIP_LAUNCH_CHECK();
}

// y (=|+=) sum c_k X_k for any number of vectors (batches of kMaxComb).
template <class T, class TY>
void launch_combine(Ctx& c, int64_t N, const std::vector<const T*>& v,
                    const std::vector<double>& coef, TY* y, bool accumulate) {
    size_t done = 0;
    bool acc = accumulate;
    if (v.empty()) {
        if (!accumulate) launch_fill<TY>(c, N, TY(0), y);
        return;
    }
    while (done < v.size()) {
        CombArgs<T> a{};
        a.count = (int)std::min<size_t>(kMaxComb, v.size() - done);
        bool al = aligned2(y, sizeof(TY));
        for (int k = 0; k < a.count; ++k) {
            a.v[k] = v[done + k];
            a.c[k] = coef[done + k];
            al = al && aligned2(a.v[k], sizeof(T));
        }
        if (al)
            hipLaunchKernelGGL((k_combine<true, T, TY>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                               c.stream, N, a, y, acc ? 1 : 0);
        else
            hipLaunchKernelGGL((k_combine<false, T, TY>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                               c.stream, N, a, y, acc ? 1 : 0);
        NUMA_HIP_LAUNCH_CHECK();
        done += a.count;
        acc = true;
    }
}

// ||x||^2 (one partial per block) and E = lambda e^x in FP32, one pass over
// x: the per-Newton-point data of the mixed-precision J v (EpJvFDe).
// lambda from *lam_ptr when set.
template <bool VEC>
__device__ inline void xnorm_exp_body(GridPos g, int64_t N, const double* x, double lambda,
                                      const double* lam_ptr, float* E, Reduce red) {
    const double lam = lam_ptr ? *lam_ptr : lambda;
    double acc = 0.0;
    for_each_pair(
        g, N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            Pair<double> x0{}, x1{};
            if (ok0) x0 = ldp<VEC>(x, q0);
            if (ok1) x1 = ldp<VEC>(x, q1);
            acc += x0.a * x0.a;
            acc += x0.b * x0.b;
            acc += x1.a * x1.a;
            acc += x1.b * x1.b;
            if (ok0) stp<true>(E, q0, Pair<float>{(float)(lam * exp(x0.a)), (float)(lam * exp(x0.b))});
            if (ok1) stp<true>(E, q1, Pair<float>{(float)(lam * exp(x1.a)), (float)(lam * exp(x1.b))});
        },
        [&](int64_t i) {
            acc += x[i] * x[i];
            E[i] = (float)(lam * exp(x[i]));
        });
    block_partial<kBlock>(acc, red, g.bid);
}

}  // namespace numa_hip
