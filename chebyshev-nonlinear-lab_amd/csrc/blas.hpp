// blas.hpp - BLAS-1 and Arnoldi kernels (operators/operator.hpp:19-41,
// linalg/gmres.hpp:141-153) on gfx950.
//
// Grid-stride streaming kernels with a fixed launch shape per vector length,
// so every reduction sums the same partials in the same order (bitwise
// repeatable).  Dots accumulate in FP64 whatever the storage type.
#pragma once

#include "common.hpp"

namespace numa_hip {

template <class T>
__device__ inline double dv(const T* p, int64_t i) {
    return static_cast<double>(p[i]);
}

// sum_i a_i b_i (b may equal a) -> epilogue.
template <class TA, class TB>
__global__ __launch_bounds__(kBlock) void k_dot(int64_t N, const TA* a, const TB* b, Reduce red,
                                                 Epilogue ep) {
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        acc += dv(a, i) * dv(b, i);
    reduce_finish<kBlock>(block_sum<kBlock>(acc), red, ep, gridDim.x, blockIdx.x);
}

template <class TA, class TB>
void launch_dot(Ctx& c, int64_t N, const TA* a, const TB* b, const Epilogue& ep) {
    hipLaunchKernelGGL((k_dot<TA, TB>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N, a, b,
                       c.red, ep);
    NUMA_HIP_LAUNCH_CHECK();
}

// y += a x  (operator.hpp:32-41)
template <class TX, class TY>
__global__ __launch_bounds__(kBlock) void k_axpy(int64_t N, double a, const TX* x, TY* y) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        y[i] = static_cast<TY>(dv(y, i) + a * dv(x, i));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_fill(int64_t N, T v, T* y) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        y[i] = v;
}

template <class TI, class TO>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t N, const TI* x, TO* y) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        y[i] = static_cast<TO>(x[i]);
}

// One fused modified-Gram-Schmidt pass (gmres.hpp:141-144):
//   w -= h_i * (s_i V_i)            with h_i = *hi (already the true dot)
//   next:  sum w * V_{i+1}   -> H[i+1][j] = s_{i+1} * sum   (EPI_STORE, mul_ptr)
//   last:  sum w * w         -> H[j+1][j] = sqrt, s_{j+1} = 1/sqrt (EPI_NORM_SCALE)
// V_i are stored unnormalised with scales s_i.
template <class T>
__global__ __launch_bounds__(kBlock) void k_mgs(int64_t N, T* w, const T* Vi, const double* hi,
                                                 const double* si, const T* Vnext, Reduce red,
                                                 Epilogue ep) {
#pragma clang fp contract(off)
    const double coef = (*hi) * (*si);
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock) {
        const T wn = static_cast<T>(dv(w, i) - coef * dv(Vi, i));
        w[i] = wn;
        const double wd = static_cast<double>(wn);
        acc += Vnext ? wd * dv(Vnext, i) : wd * wd;
    }
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

template <class T, class TY>
__global__ __launch_bounds__(kBlock) void k_combine(int64_t N, CombArgs<T> a, TY* y,
                                                     int accumulate) {
#pragma clang fp contract(off)
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock) {
        double s = accumulate ? dv(y, i) : 0.0;
        for (int k = 0; k < a.count; ++k) s += a.c[k] * dv(a.v[k], i);
        y[i] = static_cast<TY>(s);
    }
}

// z = r / d   (BratuJacobi::Bound, bratu3d.hpp:245-248)
template <class T>
__global__ __launch_bounds__(kBlock) void k_diag_solve(int64_t N, const T* r, const double* d,
                                                        T* z) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * kBlock)
        z[i] = static_cast<T>(dv(r, i) / d[i]);
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
    hipLaunchKernelGGL((k_fill<T>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N, v, y);
    NUMA_HIP_LAUNCH_CHECK();
}

template <class TI, class TO>
void launch_convert(Ctx& c, int64_t N, const TI* x, TO* y) {
    hipLaunchKernelGGL((k_convert<TI, TO>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N,
                       x, y);
    NUMA_HIP_LAUNCH_CHECK();
}

template <class TX, class TY>
void launch_axpy(Ctx& c, int64_t N, double a, const TX* x, TY* y) {
    hipLaunchKernelGGL((k_axpy<TX, TY>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream, N, a,
                       x, y);
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
        for (int k = 0; k < a.count; ++k) {
            a.v[k] = v[done + k];
            a.c[k] = coef[done + k];
        }
        hipLaunchKernelGGL((k_combine<T, TY>), dim3(stream_blocks(N)), dim3(kBlock), 0, c.stream,
                           N, a, y, acc ? 1 : 0);
        NUMA_HIP_LAUNCH_CHECK();
        done += a.count;
        acc = true;
    }
}

}  // namespace numa_hip
