// chandra_gemm.hpp - MFMA contractions Y = K X of the Chandrasekhar
// H-equation operator across a batch of instances (gfx950).
//
//   K_ij = 0.5 * w_j * mu_i / (mu_i + mu_j)      (problems/chandrasekhar_op.hpp:121-123
//                                                 without the factor c)
//   Y[b][i] = sum_j K_ij X[b][j]                 (the factor c_b is applied by the caller)
//
// Two forms of each contraction: the generating kernels (k_gemm_f64,
// k_gemm_bf16x3), where each wave computes the K values of its MFMA A operand
// in registers from mu and w, and the stored-K kernels (k_gemm_f64_kd,
// k_gemm_bf16x3_k), which stream K written once by k_fill_k with the same
// expressions and return the same bits.  X is instance-major ([b][j], j
// contiguous) and is staged through LDS per k-chunk, shared by the waves of a
// workgroup.  Sizes are padded: N to a multiple of 128, B to a multiple of 64;
// padded nodes have w = 0 (so they add nothing) and mu = 1.
//
// k_gemm_f64:   v_mfma_f64_16x16x4f64, FP64 throughout (F(x) and the FP64 mode).
//   Operand layout (checked on the device): A lane l = K[row l%16][k l/16],
//   B lane l = X[col l%16][k l/16], D lane l reg r = (row 4r + l/16, col l%16).
//   Within a chunk of 32 k, MFMA step s (0..7) takes k = 8 (l/16) + s from lane
//   l, so each lane's 8 values of a chunk are contiguous in k (two 16-byte loads
//   of stored K instead of eight 8-byte gathers).
// k_gemm_bf16x3: split-BF16 with FP32 accumulation: K = Kh + Kl and X = Xh + Xl
//   (bf16 parts); Y ~ Kh Xh + Kh Xl + Kl Xh with v_mfma_f32_32x32x16_bf16.
//   Operand layout (checked on the device): A lane l = K[row l%32][k 8(l/32)+e],
//   B lane l = X[col l%32][k 8(l/32)+e], D lane l reg r =
//   (row 8(r/4) + 4(l/32) + r%4, col l%32).  Split-K partial sums are reduced in
//   a fixed order by k_splitk_reduce, or by the kernel that consumes them.
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
// tiles), k chunks of 32 staged in LDS.
constexpr int kG64Rows = 64, kG64Cols = 64, kG64K = 32, kG64Stride = kG64K + 1;

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
        // stage X[b0 .. b0+64)[j0 .. j0+32): 2048 doubles, 8 per thread
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int e = threadIdx.x + t * 256;
            const int c = e >> 5, k = e & 31;
            xs[c * kG64Stride + k] = X[(size_t)(b0 + c) * Np + j0 + k];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kG64K / 4; ++s) {
            const int kk = 8 * (lane >> 4) + s;
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

// ---- stored K ----------------------------------------------------------------
// When N is small enough the handle computes K once (FP64, and for split-BF16
// its bf16 hi / lo parts), row-major [Np][Np], with exactly the expressions of
// the generating kernels above.  The stored-K kernels below issue the same
// MFMA sequence per accumulator on the same operand values, so they return the
// same bits, and stream K from memory instead of spending a division (and for
// bf16 a split) per entry.  Both prefetch the next k chunk (K into registers,
// X into the other half of a double-buffered LDS tile) while the current one
// is on the matrix cores: one barrier per chunk.  Rows are ldd (FP64) and ldb
// (bf16) elements apart.
__global__ __launch_bounds__(256) void k_fill_k(int Np, int ldd, int ldb,
                                                 const double* __restrict__ mu,
                                                 const double* __restrict__ w,
                                                 const float* __restrict__ muf,
                                                 const float* __restrict__ wf,
                                                 double* __restrict__ Kd, __bf16* __restrict__ Kh,
                                                 __bf16* __restrict__ Kl) {
    const int i = blockIdx.y;
    const double mu_i = mu[i];
    const double half_mu_i = 0.5 * mu_i;
    const float muf_i = muf[i];
    const float half_muf_i = 0.5f * muf_i;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < Np; j += gridDim.x * 256) {
        Kd[(size_t)i * ldd + j] = (half_mu_i * w[j]) / (mu_i + mu[j]);
        if (Kh) {
            const float kv = (half_muf_i * wf[j]) / (muf_i + muf[j]);
            __bf16 h, l;
            split_bf16(kv, h, l);
            Kh[(size_t)i * ldb + j] = h;
            Kl[(size_t)i * ldb + j] = l;
        }
    }
}

// FP64 from stored K.  Wave tile 16 RT rows x 16 CT instances, workgroup 4 waves
// (64 RT rows x 16 CT instances), k chunks of 32.  Two register sets of K alternate (the loop is
// unrolled by two chunks): the loads for chunk c+1 are issued, unconditionally
// (the last chunk reloads itself), before chunk c's MFMAs, which wait only for
// the older set.
constexpr int kGdK = kG64K, kGdStride = kGdK + 1;
using kI0 = std::integral_constant<int, 0>;
using kI1 = std::integral_constant<int, 1>;

template <int RT, int CT>
__global__ __launch_bounds__(256) void k_gemm_f64_kd(int Np, int ldk, const double* __restrict__ Kd,
                                                      const double* __restrict__ X,
                                                      double* __restrict__ Y) {
    constexpr int XT = 2 * CT;  // staged values per thread
    __shared__ double xs[2][16 * CT * kGdStride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = blockIdx.x * (64 * RT) + wave * 16 * RT, b0 = blockIdx.y * 16 * CT;
    const int q = lane >> 4;
    const int nch = Np / kGdK;
    const double* krow[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) krow[r] = Kd + (size_t)(i0 + r * 16 + (lane & 15)) * ldk + 8 * q;
    // staging: thread -> k = tid % 32 of instances tid / 32 + 8 t, t < 2 CT
    const int sc = threadIdx.x >> 5, sk = threadIdx.x & 31;
    const double* xsrc = X + (size_t)(b0 + sc) * Np + sk;
    double xr[XT], a[2][RT][8];
    auto load = [&](int c, double(&ak)[RT][8]) {
        const int j = c * kGdK;
#pragma unroll
        for (int t = 0; t < XT; ++t) xr[t] = xsrc[(size_t)(8 * t) * Np + j];
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int s = 0; s < 8; s += 2) {
                const double2 v2 = *reinterpret_cast<const double2*>(krow[r] + j + s);
                ak[r][s] = v2.x;
                ak[r][s + 1] = v2.y;
            }
    };
    auto stage = [&](double* buf) {
#pragma unroll
        for (int t = 0; t < XT; ++t) buf[(sc + 8 * t) * kGdStride + sk] = xr[t];
    };
    doublex4 acc[RT][CT] = {};
    auto step = [&](auto pc, int c) {
        constexpr int p = decltype(pc)::value;
        load(c + 1 < nch ? c + 1 : c, a[p ^ 1]);
        const double* xb = xs[p] + (lane & 15) * kGdStride + 8 * q;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                const double bv = xb[t * 16 * kGdStride + s];
#pragma unroll
                for (int r = 0; r < RT; ++r)
                    acc[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[p][r][s], bv, acc[r][t], 0, 0, 0);
            }
        }
        stage(xs[p ^ 1]);
        __syncthreads();
    };
    load(0, a[0]);
    stage(xs[0]);
    __syncthreads();
    int c = 0;
    for (; c + 1 < nch; c += 2) {
        step(kI0{}, c);
        step(kI1{}, c + 1);
    }
    if (c < nch) step(kI0{}, c);
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int t = 0; t < CT; ++t) {
            const int col = b0 + t * 16 + (lane & 15);
#pragma unroll
            for (int rr = 0; rr < 4; ++rr)
                Y[(size_t)col * Np + i0 + r * 16 + 4 * rr + q] = acc[r][t][rr];
        }
}

// k_gemm_f64_kd with the operand loads PF chunks ahead of the MFMAs (a ring
// of PF register sets of K and of X values, the X slot of chunk c + 1 staged
// to LDS while chunk c is contracted).  For narrow blocks (16 instances a
// wave), where a wave's MFMAs per chunk are few and one chunk in flight
// leaves it waiting on memory.  Every output element gets the same MFMA
// steps in the same order as in k_gemm_f64_kd.  Needs Np % (32 PF) == 0.
template <int RT, int CT, int PF>
__global__ __launch_bounds__(256) void k_gemm_f64_kd_pf(int Np, int ldk,
                                                         const double* __restrict__ Kd,
                                                         const double* __restrict__ X,
                                                         double* __restrict__ Y) {
    constexpr int XT = 2 * CT;
    __shared__ double xs[2][16 * CT * kGdStride];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = blockIdx.x * (64 * RT) + wave * 16 * RT, b0 = blockIdx.y * 16 * CT;
    const int q = lane >> 4;
    const int nch = Np / kGdK;
    const double* krow[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) krow[r] = Kd + (size_t)(i0 + r * 16 + (lane & 15)) * ldk + 8 * q;
    const int sc = threadIdx.x >> 5, sk = threadIdx.x & 31;
    const double* xsrc = X + (size_t)(b0 + sc) * Np + sk;
    double xr[PF][XT], a[PF][RT][8];
    // chunk c (the last chunk again past the end) into ring slot S
    auto load = [&](int c, auto slot) {
        constexpr int S = decltype(slot)::value;
        const int j = (c < nch ? c : nch - 1) * kGdK;
#pragma unroll
        for (int t = 0; t < XT; ++t) xr[S][t] = xsrc[(size_t)(8 * t) * Np + j];
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int s = 0; s < 8; s += 2) {
                const double2 v2 = *reinterpret_cast<const double2*>(krow[r] + j + s);
                a[S][r][s] = v2.x;
                a[S][r][s + 1] = v2.y;
            }
    };
    auto stage = [&](double* buf, auto slot) {
        constexpr int S = decltype(slot)::value;
#pragma unroll
        for (int t = 0; t < XT; ++t) buf[(sc + 8 * t) * kGdStride + sk] = xr[S][t];
    };
    doublex4 acc[RT][CT] = {};
    // chunk c in slot S = c % PF, LDS buffer c & 1 (PF even)
    auto step = [&](auto slot, int c) {
        constexpr int S = decltype(slot)::value;
        constexpr int N1 = (S + 1) % PF;
        const double* xb = xs[S & 1] + (lane & 15) * kGdStride + 8 * q;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                const double bv = xb[t * 16 * kGdStride + s];
#pragma unroll
                for (int r = 0; r < RT; ++r)
                    acc[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[S][r][s], bv, acc[r][t], 0, 0, 0);
            }
        }
        stage(xs[(S + 1) & 1], std::integral_constant<int, N1>{});
        load(c + PF, slot);
        __syncthreads();
    };
    static_for<PF>([&](auto sl) { load(decltype(sl)::value, sl); });
    stage(xs[0], std::integral_constant<int, 0>{});
    __syncthreads();
    for (int c = 0; c < nch; c += PF)
        static_for<PF>([&](auto sl) { step(sl, c + decltype(sl)::value); });
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int t = 0; t < CT; ++t) {
            const int col = b0 + t * 16 + (lane & 15);
#pragma unroll
            for (int rr = 0; rr < 4; ++rr)
                Y[(size_t)col * Np + i0 + r * 16 + 4 * rr + q] = acc[r][t][rr];
        }
}

// Split-BF16 from stored Kh, Kl.  WR waves per workgroup, wave tile 32 RT rows x
// 32 CT instances (WR 32 RT rows per workgroup), k chunks of KC (32, 64 or 128,
// dividing the split-K slice); blockIdx.z = split-K slice, as in k_gemm_bf16x3.
// Same register double-buffering as k_gemm_f64_kd.  LDS row stride KC + 8 bf16
// (a multiple of 4 dwords, odd multiple of 4 modulo 32) keeps the 16-byte
// operand reads of 16 consecutive lanes on distinct banks.
template <int WR, int RT, int CT, int KC>
__global__ __launch_bounds__(WR * 64) void k_gemm_bf16x3_k(int Np, int ldk, int ksplit,
                                                            const __bf16* __restrict__ Kh,
                                                            const __bf16* __restrict__ Kl,
                                                            const __bf16* __restrict__ Xh,
                                                            const __bf16* __restrict__ Xl,
                                                            float* __restrict__ P) {
    constexpr int ST = KC + 8;                       // LDS row stride (bf16)
    constexpr int NS = KC / 16;                      // MFMA k steps per chunk
    constexpr int CPR = KC / 8;                      // 16-byte pieces per instance row
    constexpr int NC = 32 * CT;                      // instances per workgroup
    constexpr int NP = NC * CPR;                     // staging pieces per part
    constexpr int SC = (NP + WR * 64 - 1) / (WR * 64);  // per thread
    static_assert(NP % (WR * 64) == 0 || NP < WR * 64, "staging must cover the tile");
    __shared__ __attribute__((aligned(16))) __bf16 xh[2][NC * ST];
    __shared__ __attribute__((aligned(16))) __bf16 xl[2][NC * ST];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = blockIdx.x * (WR * 32 * RT) + wave * 32 * RT, b0 = blockIdx.y * NC;
    const int klen = Np / ksplit;
    const int kbeg = blockIdx.z * klen;
    const int nch = klen / KC;
    const int kq = (lane >> 5) * 8;
    const __bf16* kh[RT];
    const __bf16* kl[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const size_t o = (size_t)(i0 + r * 32 + (lane & 31)) * ldk + kbeg + kq;
        kh[r] = Kh + o;
        kl[r] = Kl + o;
    }
    // staging: piece e = tid + u * WR * 64 -> instance e / CPR, k 8 (e % CPR)
    size_t xo[SC];
    int lo[SC];
#pragma unroll
    for (int u = 0; u < SC; ++u) {
        const int e = threadIdx.x + u * WR * 64;
        xo[u] = (size_t)(b0 + e / CPR) * Np + kbeg + (e % CPR) * 8;
        lo[u] = (e / CPR) * ST + (e % CPR) * 8;
    }
    bf16x8 rh[SC], rl[SC], ah[2][RT][NS], al[2][RT][NS];
    auto load = [&](int c, bf16x8(&nh)[RT][NS], bf16x8(&nl)[RT][NS]) {
        const int j = c * KC;
#pragma unroll
        for (int u = 0; u < SC; ++u) {
            if (NP >= WR * 64 || (int)threadIdx.x < NP) {
                rh[u] = *reinterpret_cast<const bf16x8*>(&Xh[xo[u] + j]);
                rl[u] = *reinterpret_cast<const bf16x8*>(&Xl[xo[u] + j]);
            }
        }
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                nh[r][s] = *reinterpret_cast<const bf16x8*>(kh[r] + j + s * 16);
                nl[r][s] = *reinterpret_cast<const bf16x8*>(kl[r] + j + s * 16);
            }
    };
    auto stage = [&](__bf16* bh, __bf16* bl) {
#pragma unroll
        for (int u = 0; u < SC; ++u) {
            if (NP >= WR * 64 || (int)threadIdx.x < NP) {
                *reinterpret_cast<bf16x8*>(&bh[lo[u]]) = rh[u];
                *reinterpret_cast<bf16x8*>(&bl[lo[u]]) = rl[u];
            }
        }
    };
    floatx16 acc[RT][CT] = {};
    auto step = [&](auto pc, int c) {
        constexpr int p = decltype(pc)::value;
        load(c + 1 < nch ? c + 1 : c, ah[p ^ 1], al[p ^ 1]);
        // keep the next chunk's loads ahead of this chunk's MFMAs (the
        // scheduler otherwise sinks them to the end, under register pressure)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
#pragma unroll
            for (int t = 0; t < CT; ++t) {
                const int o = (t * 32 + (lane & 31)) * ST + s * 16 + kq;
                const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&xh[p][o]);
                const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&xl[p][o]);
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    acc[r][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[p][r][s], bh, acc[r][t], 0, 0, 0);
                    acc[r][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[p][r][s], bl, acc[r][t], 0, 0, 0);
                    acc[r][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[p][r][s], bh, acc[r][t], 0, 0, 0);
                }
            }
        }
        stage(xh[p ^ 1], xl[p ^ 1]);
        __syncthreads();
    };
    load(0, ah[0], al[0]);
    stage(xh[0], xl[0]);
    __syncthreads();
    int c = 0;
    for (; c + 1 < nch; c += 2) {
        step(kI0{}, c);
        step(kI1{}, c + 1);
    }
    if (c < nch) step(kI0{}, c);
    const int Bp = gridDim.y * NC;
    float* out = P + (size_t)blockIdx.z * Bp * Np;
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int t = 0; t < CT; ++t) {
            const int col = b0 + t * 32 + (lane & 31);
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) {
                const int orow = i0 + r * 32 + 8 * (rr >> 2) + 4 * (lane >> 5) + (rr & 3);
                out[(size_t)col * Np + orow] = acc[r][t][rr];
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
