// chandra_gemm.hpp - MFMA contractions Y = K X of the Chandrasekhar
// H-equation operator across a batch of instances (gfx950).
//
//   K_ij = 0.5 * w_j * mu_i / (mu_i + mu_j)      (problems/chandrasekhar_op.hpp:121-123
//                                                 without the factor c)
//   Y[b][i] = sum_j K_ij X[b][j]                 (the factor c_b is applied by the caller)
//
// K (N x N) is never stored: each wave generates the K values of its MFMA A
// operand in registers from mu and w.  X is instance-major ([b][j], j
// contiguous) and is staged through LDS per k-chunk, shared by the waves of a
// workgroup.  Sizes are padded: N to a multiple of 128, B to a multiple of 64;
// padded nodes have w = 0 (so they add nothing) and mu = 1.
//
// k_gemm_f64:   v_mfma_f64_16x16x4f64, FP64 throughout (F(x) and the FP64 mode).
//   Operand layout (checked on the device): A lane l = K[row l%16][k l/16],
//   B lane l = X[col l%16][k l/16], D lane l reg r = (row 4r + l/16, col l%16).
// k_gemm_bf16x3: split-BF16 with FP32 accumulation: K = Kh + Kl and X = Xh + Xl
//   (bf16 parts); Y ~ Kh Xh + Kh Xl + Kl Xh with v_mfma_f32_32x32x16_bf16.
//   Operand layout (checked on the device): A lane l = K[row l%32][k 8(l/32)+e],
//   B lane l = X[col l%32][k 8(l/32)+e], D lane l reg r =
//   (row 8(r/4) + 4(l/32) + r%4, col l%32).  Split-K partial sums are reduced in
//   a fixed order by k_splitk_reduce.
#pragma once

#include "common.hpp"

namespace numa_hip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kChPadN = 128;  // N padding (rows per bf16 workgroup)
constexpr int kChPadB = 64;   // batch padding (columns per workgroup)

// ---- FP64 -----------------------------------------------------------------
// Workgroup: 4 waves, 64 rows (16 per wave) x 64 instances (4 MFMA column
// tiles), k chunks of 16 staged in LDS.
constexpr int kG64Rows = 64, kG64Cols = 64, kG64K = 16, kG64Stride = kG64K + 1;

__global__ __launch_bounds__(256) void k_gemm_f64(int Np, const double* __restrict__ mu,
                                                   const double* __restrict__ w,
                                                   const double* __restrict__ X,
                                                   double* __restrict__ Y) {
    __shared__ double xs[kG64Cols * kG64Stride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = blockIdx.x * kG64Rows, b0 = blockIdx.y * kG64Cols;
    const int row = i0 + wave * 16 + (lane & 15);
    const double mu_i = mu[row];
    const double half_mu_i = 0.5 * mu_i;
    doublex4 acc[4] = {};
    for (int j0 = 0; j0 < Np; j0 += kG64K) {
        // stage X[b0 .. b0+64)[j0 .. j0+16): 1024 doubles, 4 per thread
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int e = threadIdx.x + t * 256;
            const int c = e >> 4, k = e & 15;
            xs[c * kG64Stride + k] = X[(size_t)(b0 + c) * Np + j0 + k];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kG64K / 4; ++s) {
            const int kk = s * 4 + (lane >> 4);
            const int j = j0 + kk;
            const double a = (half_mu_i * w[j]) / (mu_i + mu[j]);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const double bv = xs[(t * 16 + (lane & 15)) * kG64Stride + kk];
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[t], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int col = b0 + t * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int orow = i0 + wave * 16 + 4 * r + (lane >> 4);
            Y[(size_t)col * Np + orow] = acc[t][r];
        }
    }
}

// ---- split BF16 ------------------------------------------------------------
// Workgroup: 4 waves, 128 rows (32 per wave) x 64 instances (2 MFMA column
// tiles), k chunks of 32 staged in LDS (row stride 40 bf16 = 80 B, which makes
// the 16-byte operand reads bank-conflict free).  blockIdx.z = split-K slice.
constexpr int kGbRows = 128, kGbCols = 64, kGbK = 32, kGbStride = 40;

__device__ inline void split_bf16(float v, __bf16& hi, __bf16& lo) {
    hi = static_cast<__bf16>(v);
    lo = static_cast<__bf16>(v - static_cast<float>(hi));
}

__global__ __launch_bounds__(256) void k_gemm_bf16x3(int Np, int ksplit,
                                                      const float* __restrict__ mu,
                                                      const float* __restrict__ w,
                                                      const __bf16* __restrict__ Xh,
                                                      const __bf16* __restrict__ Xl,
                                                      float* __restrict__ P) {
    __shared__ __attribute__((aligned(16))) __bf16 xh[kGbCols * kGbStride];
    __shared__ __attribute__((aligned(16))) __bf16 xl[kGbCols * kGbStride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = blockIdx.x * kGbRows, b0 = blockIdx.y * kGbCols;
    const int klen = Np / ksplit;
    const int kbeg = blockIdx.z * klen, kend = kbeg + klen;
    const int row = i0 + wave * 32 + (lane & 31);
    const float mu_i = mu[row];
    const float half_mu_i = 0.5f * mu_i;
    const int kq = (lane >> 5) * 8;
    floatx16 acc[2] = {};
    for (int j0 = kbeg; j0 < kend; j0 += kGbK) {
        // stage Xh, Xl [b0 .. b0+64)[j0 .. j0+32): 64 x 32 bf16 each = 256
        // chunks of 8; one 16-byte chunk per thread per part
        {
            const int c = threadIdx.x >> 2, q = (threadIdx.x & 3) * 8;
            const size_t g = (size_t)(b0 + c) * Np + j0 + q;
            *reinterpret_cast<bf16x8*>(&xh[c * kGbStride + q]) =
                *reinterpret_cast<const bf16x8*>(&Xh[g]);
            *reinterpret_cast<bf16x8*>(&xl[c * kGbStride + q]) =
                *reinterpret_cast<const bf16x8*>(&Xl[g]);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kGbK / 16; ++s) {
            const int jb = j0 + s * 16 + kq;
            bf16x8 ah, al;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float kv = (half_mu_i * w[jb + e]) / (mu_i + mu[jb + e]);
                __bf16 h, l;
                split_bf16(kv, h, l);
                ah[e] = h;
                al[e] = l;
            }
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int c = t * 32 + (lane & 31);
                const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&xh[c * kGbStride + s * 16 + kq]);
                const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&xl[c * kGbStride + s * 16 + kq]);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[t], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int Bp = gridDim.y * kGbCols;
    float* out = P + (size_t)blockIdx.z * Bp * Np;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = b0 + t * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int orow = i0 + wave * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            out[(size_t)col * Np + orow] = acc[t][r];
        }
    }
}

// X (FP64) -> bf16 hi / lo parts, over the padded [Bp][Np] array.
__global__ __launch_bounds__(256) void k_split_x(int64_t total, const double* __restrict__ X,
                                                  __bf16* __restrict__ Xh,
                                                  __bf16* __restrict__ Xl) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * 256) {
        __bf16 h, l;
        split_bf16(static_cast<float>(X[i]), h, l);
        Xh[i] = h;
        Xl[i] = l;
    }
}

// Y[b][i] = sum over slices of P[s][b][i] (fixed order), FP64 output.
__global__ __launch_bounds__(256) void k_splitk_reduce(int64_t total, int ksplit,
                                                        const float* __restrict__ P,
                                                        double* __restrict__ Y) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * 256) {
        double s = 0.0;
        for (int k = 0; k < ksplit; ++k) s += (double)P[(size_t)k * total + i];
        Y[i] = s;
    }
}

}  // namespace numa_hip
