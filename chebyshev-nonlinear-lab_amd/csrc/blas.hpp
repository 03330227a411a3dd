// blas.hpp - BLAS-1 and Arnoldi kernels (operators/operator.hpp:19-41,
// linalg/gmres.hpp:141-153) on gfx950.
//
// Grid-stride streaming kernels: each thread handles 4 consecutive elements
// per iteration with one 16-byte (FP32) or two 16-byte (FP64) vector loads
// per stream, so enough bytes are in flight per wave to keep HBM busy.  The
// launch shape is a fixed function of the length, so every reduction sums the
// same partials in the same order (bitwise repeatable).  Dots accumulate in
// FP64 whatever the storage type.  Pointers that are not 4-element aligned
// take the scalar-load variant of the same kernel.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace numa_hip {

template <class T>
__device__ inline double dv(const T* p, int64_t i) {
    return static_cast<double>(p[i]);
}

template <class T>
struct alignas(4 * sizeof(T)) V4 {
    T v[4];
};

template <bool VEC, class T>
__device__ inline V4<T> ld4(const T* p, int64_t i) {
    if constexpr (VEC) {
        return *reinterpret_cast<const V4<T>*>(p + i);
    } else {
        V4<T> r;
#pragma unroll
        for (int e = 0; e < 4; ++e) r.v[e] = p[i + e];
        return r;
    }
}

template <bool VEC, class T>
__device__ inline void st4(T* p, int64_t i, const V4<T>& r) {
    if constexpr (VEC) {
        *reinterpret_cast<V4<T>*>(p + i) = r;
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) p[i + e] = r.v[e];
    }
}

// Visits [0, N) as groups of 4 (f4(base)) plus a scalar tail (f1(i)) handled
// by block 0.
template <bool VEC, class F4, class F1>
__device__ inline void for_each4(int64_t N, F4&& f4, F1&& f1) {
    const int64_t N4 = N >> 2;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < N4;
         q += (int64_t)gridDim.x * kBlock)
        f4(q << 2);
    const int64_t rem = N - (N4 << 2);
    if (blockIdx.x == 0 && threadIdx.x < rem) f1((N4 << 2) + threadIdx.x);
}

inline bool aligned4(const void* p, size_t elem) {
    return p == nullptr || (reinterpret_cast<uintptr_t>(p) % (4 * elem)) == 0;
}

// sum_i a_i b_i (b may equal a) -> epilogue.
template <bool VEC, class TA, class TB>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t N, const TA* a, const TB* b, Reduce red,
                                                 Epilogue ep) {
    double acc = 0.0;
    for_each4<VEC>(
        N,
        [&](int64_t i) {
            const V4<TA> x = ld4<VEC>(a, i);
            const V4<TB> y = ld4<VEC>(b, i);
#pragma unroll
            for (int e = 0; e < 4; ++e) acc += (double)x.v[e] * (double)y.v[e];
        },
        [&](int64_t i) { acc += dv(a, i) * dv(b, i); });
    reduce_finish<kBlock>(block_sum<kBlock>(acc), red, ep, gridDim.x, blockIdx.x);
}

template <class TA, class TB>
void launch_dot(Ctx& c, int64_t N, const TA* a, const TB* b, const Epilogue& ep) {
    if (aligned4(a, sizeof(TA)) && aligned4(b, sizeof(TB)))
        hipLaunchKernelGGL((k_dot<true, TA, TB>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, a, b, c.red, ep);
    else
        hipLaunchKernelGGL((k_dot<false, TA, TB>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, a, b, c.red, ep);
    NUMA_HIP_LAUNCH_CHECK();
}

// y += a x  (operator.hpp:32-41)
template <bool VEC, class TX, class TY>
__global__ __launch_bounds__(kBlock) void k_axpy(int64_t N, double a, const TX* x, TY* y) {
#pragma clang fp contract(off)
    for_each4<VEC>(
        N,
        [&](int64_t i) {
            const V4<TX> xv = ld4<VEC>(x, i);
            V4<TY> yv = ld4<VEC>(y, i);
#pragma unroll
            for (int e = 0; e < 4; ++e) yv.v[e] = static_cast<TY>((double)yv.v[e] + a * (double)xv.v[e]);
            st4<VEC>(y, i, yv);
        },
        [&](int64_t i) { y[i] = static_cast<TY>(dv(y, i) + a * dv(x, i)); });
}

template <bool VEC, class T>
__global__ __launch_bounds__(kBlock) void k_fill(int64_t N, T v, T* y) {
    for_each4<VEC>(
        N,
        [&](int64_t i) {
            V4<T> r;
#pragma unroll
            for (int e = 0; e < 4; ++e) r.v[e] = v;
            st4<VEC>(y, i, r);
        },
        [&](int64_t i) { y[i] = v; });
}

template <bool VEC, class TI, class TO>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t N, const TI* x, TO* y) {
    for_each4<VEC>(
        N,
        [&](int64_t i) {
            const V4<TI> a = ld4<VEC>(x, i);
            V4<TO> r;
#pragma unroll
            for (int e = 0; e < 4; ++e) r.v[e] = static_cast<TO>(a.v[e]);
            st4<VEC>(y, i, r);
        },
        [&](int64_t i) { y[i] = static_cast<TO>(x[i]); });
}

// One fused modified-Gram-Schmidt pass (gmres.hpp:141-144):
//   w -= h_i * (s_i V_i)            with h_i = *hi (already the true dot)
//   next:  sum w * V_{i+1}   -> H[i+1][j] = s_{i+1} * sum   (EPI_STORE, mul_ptr)
//   last:  sum w * w         -> H[j+1][j] = sqrt, s_{j+1} = 1/sqrt (EPI_NORM_SCALE)
// V_i are stored unnormalised with scales s_i.
template <bool VEC, class T>
__global__ __launch_bounds__(kBlock) void k_mgs(int64_t N, T* w, const T* Vi, const double* hi,
                                                 const double* si, const T* Vnext, Reduce red,
                                                 Epilogue ep) {
#pragma clang fp contract(off)
    const double coef = (*hi) * (*si);
    double acc = 0.0;
    for_each4<VEC>(
        N,
        [&](int64_t i) {
            V4<T> wv = ld4<VEC>(w, i);
            const V4<T> vv = ld4<VEC>(Vi, i);
#pragma unroll
            for (int e = 0; e < 4; ++e) wv.v[e] = static_cast<T>((double)wv.v[e] - coef * (double)vv.v[e]);
            st4<VEC>(w, i, wv);
            if (Vnext) {
                const V4<T> nv = ld4<VEC>(Vnext, i);
#pragma unroll
                for (int e = 0; e < 4; ++e) acc += (double)wv.v[e] * (double)nv.v[e];
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) acc += (double)wv.v[e] * (double)wv.v[e];
            }
        },
        [&](int64_t i) {
            const T wn = static_cast<T>(dv(w, i) - coef * dv(Vi, i));
            w[i] = wn;
            const double wd = static_cast<double>(wn);
            acc += Vnext ? wd * dv(Vnext, i) : wd * wd;
        });
    reduce_finish<kBlock>(block_sum<kBlock>(acc), red, ep, gridDim.x, blockIdx.x);
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
__global__ __launch_bounds__(kBlock) void k_combine(int64_t N, CombArgs<T> a, TY* y,
                                                     int accumulate) {
#pragma clang fp contract(off)
    for_each4<VEC>(
        N,
        [&](int64_t i) {
            double s[4] = {0.0, 0.0, 0.0, 0.0};
            if (accumulate) {
                const V4<TY> yv = ld4<VEC>(y, i);
#pragma unroll
                for (int e = 0; e < 4; ++e) s[e] = (double)yv.v[e];
            }
            for (int k = 0; k < a.count; ++k) {
                const V4<T> xv = ld4<VEC>(a.v[k], i);
#pragma unroll
                for (int e = 0; e < 4; ++e) s[e] += a.c[k] * (double)xv.v[e];
            }
            V4<TY> r;
#pragma unroll
            for (int e = 0; e < 4; ++e) r.v[e] = static_cast<TY>(s[e]);
            st4<VEC>(y, i, r);
        },
        [&](int64_t i) {
            double s = accumulate ? dv(y, i) : 0.0;
            for (int k = 0; k < a.count; ++k) s += a.c[k] * dv(a.v[k], i);
            y[i] = static_cast<TY>(s);
        });
}

// z = r / d   (BratuJacobi::Bound, bratu3d.hpp:245-248)
template <bool VEC, class T>
__global__ __launch_bounds__(kBlock) void k_diag_solve(int64_t N, const T* r, const double* d,
                                                        T* z) {
    for_each4<VEC>(
        N,
        [&](int64_t i) {
            const V4<T> rv = ld4<VEC>(r, i);
            const V4<double> dd = ld4<VEC>(d, i);
            V4<T> o;
#pragma unroll
            for (int e = 0; e < 4; ++e) o.v[e] = static_cast<T>((double)rv.v[e] / dd.v[e]);
            st4<VEC>(z, i, o);
        },
        [&](int64_t i) { z[i] = static_cast<T>(dv(r, i) / d[i]); });
}

namespace {
// d = 6 h2inv - lambda e^u   (bratu3d.hpp:59-69)
__global__ __launch_bounds__(kBlock) void k_jacobi_diag(int64_t N, double h2inv, double lambda,
                                                         const double* u, double* d) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        d[i] = 6.0 * h2inv - lambda * exp(u[i]);
}

// out = -lambda e^u v v   (bratu3d.hpp:46-56)
__global__ __launch_bounds__(kBlock) void k_d2_analytic(int64_t N, double lambda, const double* u,
                                                         const double* v, double* out) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        out[i] = -lambda * exp(u[i]) * v[i] * v[i];
}
}  // namespace

template <class T>
void launch_fill(Ctx& c, int64_t N, T v, T* y) {
    if (aligned4(y, sizeof(T)))
        hipLaunchKernelGGL((k_fill<true, T>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N,
                           v, y);
    else
        hipLaunchKernelGGL((k_fill<false, T>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream,
                           N, v, y);
    NUMA_HIP_LAUNCH_CHECK();
}

template <class TI, class TO>
void launch_convert(Ctx& c, int64_t N, const TI* x, TO* y) {
    if (aligned4(x, sizeof(TI)) && aligned4(y, sizeof(TO)))
        hipLaunchKernelGGL((k_convert<true, TI, TO>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, x, y);
    else
        hipLaunchKernelGGL((k_convert<false, TI, TO>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, x, y);
    NUMA_HIP_LAUNCH_CHECK();
}

template <class TX, class TY>
void launch_axpy(Ctx& c, int64_t N, double a, const TX* x, TY* y) {
    if (aligned4(x, sizeof(TX)) && aligned4(y, sizeof(TY)))
        hipLaunchKernelGGL((k_axpy<true, TX, TY>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, a, x, y);
    else
        hipLaunchKernelGGL((k_axpy<false, TX, TY>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                           c.stream, N, a, x, y);
    NUMA_HIP_LAUNCH_CHECK();
}

template <class T>
void launch_mgs(Ctx& c, int64_t N, T* w, const T* Vi, const double* hi, const double* si,
                const T* Vnext, const Epilogue& ep) {
    if (aligned4(w, sizeof(T)) && aligned4(Vi, sizeof(T)) && aligned4(Vnext, sizeof(T)))
        hipLaunchKernelGGL((k_mgs<true, T>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N,
                           w, Vi, hi, si, Vnext, c.red, ep);
    else
        hipLaunchKernelGGL((k_mgs<false, T>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N,
                           w, Vi, hi, si, Vnext, c.red, ep);
    NUMA_HIP_LAUNCH_CHECK();
}

template <class T>
void launch_diag_solve(Ctx& c, int64_t N, const T* r, const double* d, T* z) {
    if (aligned4(r, sizeof(T)) && aligned4(d, sizeof(double)) && aligned4(z, sizeof(T)))
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
        bool al = aligned4(y, sizeof(TY));
        for (int k = 0; k < a.count; ++k) {
            a.v[k] = v[done + k];
            a.c[k] = coef[done + k];
            al = al && aligned4(a.v[k], sizeof(T));
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

}  // namespace numa_hip
