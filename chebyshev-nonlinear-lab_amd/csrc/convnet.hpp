// convnet.hpp - kernels of the learned preconditioner: a 3D conv net of six
// 3x3x3 layers (channels 1 -> 24 -> 24 -> 24 -> 24 -> 24 -> 1, GELU between
// layers, zero padding = the zero Dirichlet boundary) on the n^3 interior grid.
//
// Activations are channel-last bf16: act[b][p][24], p = (i*n + j)*n + k (the
// grid's own ordering), 48 bytes per point = three 16-byte chunks of 8
// channels.  The four 24 -> 24 layers (98% of the FLOPs) run on MFMA
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation) as implicit GEMMs; the first
// and last layers (1 and 24 channels on one side) run on the VALU in FP32.
//
// Parameters (FP32 master copy, flat): per layer W[Cout][Cin][3][3][3] then
// b[Cout]; W[co][ci][tap] multiplies a[p + d(tap)][ci] with
// tap = (di+1)*9 + (dj+1)*3 + (dk+1), d(tap) = (di, dj, dk).
//
// Work decomposition: a tile is 2 (i) x 8 (j) lines of 32 points along k.
// Persistent workgroups walk the tiles; worker ids are laid out so that the
// tiles processed at the same time by one XCD are contiguous.  The input halo
// of a tile, 4 x 10 x 34 points, is staged in LDS.
//
// MFMA operand layout (v_mfma_f32_32x32x16_bf16, checked on the device):
//   A lane l = A[row l%32][k 8(l/32) .. +8], B lane l = B[k 8(l/32) .. +8][col l%32],
//   D lane l reg r = D[row 8(r/4) + 4(l/32) + r%4][col l%32].
#pragma once

#include "common.hpp"

namespace numa_hip {
namespace cnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kC = 24;                      // hidden channels
constexpr int kTaps = 27;
constexpr int kGroups = kTaps * 3;          // K groups of 8: (tap, channel group)
constexpr int kSteps = (kGroups + 1) / 2;   // 41 MFMA k-steps of 16
constexpr int kTI = 2, kTJ = 8, kTK = 32;   // tile: 2 x 8 lines of 32 points
constexpr int kHI = kTI + 2, kHJ = kTJ + 2, kHK = kTK + 2;
constexpr int kHalo = kHI * kHJ * kHK;      // 1360 halo points
constexpr int kLines = kTI * kTJ;           // 16
constexpr int kTilePts = kLines * kTK;      // 512
constexpr int kWorkers = 512;               // persistent workgroups (multiple of 8)

// flat parameter layout
constexpr int64_t kW0 = 0, kB0 = kW0 + kC * kTaps;                 // first layer
constexpr int64_t kMidW = kC * kC * kTaps;                          // 15552
constexpr int64_t kMid0 = kB0 + kC;                                 // first mid layer
constexpr int64_t kMidStride = kMidW + kC;
__host__ __device__ constexpr int64_t mid_w(int l) { return kMid0 + l * kMidStride; }  // l = 0..3
__host__ __device__ constexpr int64_t mid_b(int l) { return mid_w(l) + kMidW; }
constexpr int64_t kW5 = kMid0 + 4 * kMidStride, kB5 = kW5 + kC * kTaps;
constexpr int64_t kParams = kB5 + 1;                                // 63625

// GELU(x) = x Phi(x) with the normal distribution function
//   Phi(x) = 1 - erfc(z)/2 (x >= 0),  erfc(z)/2 (x < 0),   z = |x| / sqrt 2,
// and erfc(z) = (a1 t + a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5) exp(-z^2),
// t = 1 / (1 + p z): Abramowitz & Stegun, Handbook of Mathematical Functions,
// formula 7.1.26 (absolute error of erf at most 1.5e-7, far below the bf16
// rounding of the activations).  A dozen VALU operations and two
// transcendentals (reciprocal, 2^x) per value: the epilogues apply it to 24
// channels of every point, where a library erf dominated the layer's time.
// Also returns e = exp(-x^2 / 2), the unnormalised density (for GELU').
__device__ inline float norm_cdf(float x, float& e) {
    constexpr float p = 0.3275911f, a1 = 0.254829592f, a2 = -0.284496736f,
                    a3 = 1.421413741f, a4 = -1.453152027f, a5 = 1.061405429f;
    const float z = fabsf(x) * 0.70710678118654752f;
    const float t = __builtin_amdgcn_rcpf(fmaf(p, z, 1.0f));
    const float poly = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, a5, a4), a3), a2), a1);
    e = __builtin_amdgcn_exp2f(-(z * z) * 1.4426950408889634f);  // exp(-z^2)
    const float half_erfc = 0.5f * poly * e;
    return x >= 0.0f ? 1.0f - half_erfc : half_erfc;
}
__device__ inline float gelu_f(float x) {
    float e;
    return x * norm_cdf(x, e);
}
__device__ inline float gelu_grad_f(float x) {  // Phi(x) + x phi(x)
    float e;
    const float c = norm_cdf(x, e);
    return fmaf(x * 0.39894228040143268f, e, c);
}

// Halo-point offset of a tap in a (kHI, kHJ, kHK) tile, and the 16-byte chunk
// offset of K group g = 3*tap + channel group (the missing 82nd group reuses
// the last valid address; its weights are zero).
__host__ __device__ constexpr int tap_off(int tap) {
    return (tap / 9 - 1) * kHJ * kHK + ((tap / 3) % 3 - 1) * kHK + (tap % 3 - 1);
}
__host__ __device__ constexpr int group_off(int g) {
    return g >= kGroups ? tap_off(kTaps - 1) * 3 + 2 : tap_off(g / 3) * 3 + g % 3;
}

// Smallest group offset, and the step between the two K groups 2s, 2s + 1
// that the two half-waves read at MFMA step s: one 16-byte chunk within a
// tap, or across a k-neighbour tap; 94 / 808 chunks where the pair straddles
// the end of a k row / j row of taps; 0 for the padded 82nd group.
constexpr int group_off_min() {
    int m = 0;
    for (int g = 0; g <= kGroups; ++g) m = group_off(g) < m ? group_off(g) : m;
    return m;
}
constexpr int kGroupOffMin = group_off_min();
constexpr int pair_step(int s) { return group_off(2 * s + 1) - group_off(2 * s); }
constexpr bool pair_steps_known() {
    for (int s = 0; s < kSteps; ++s) {
        const int d = pair_step(s);
        if (d != 0 && d != 1 && d != 94 && d != 808) return false;
    }
    return true;
}
static_assert(pair_steps_known(), "unexpected K group layout");

template <int TI = kTI>
struct Tiles {
    int n, nti, ntj, ntk;
    int64_t per_sample, total;
    __host__ __device__ Tiles(int n_, int ns)
        : n(n_), nti((n_ + TI - 1) / TI), ntj((n_ + kTJ - 1) / kTJ), ntk((n_ + kTK - 1) / kTK) {
        per_sample = (int64_t)nti * ntj * ntk;
        total = per_sample * ns;
    }
    __device__ void decode(int64_t t, int& b, int& i0, int& j0, int& k0) const {
        b = (int)(t / per_sample);
        int64_t r = t - (int64_t)b * per_sample;
        k0 = (int)(r % ntk) * kTK;
        r /= ntk;
        j0 = (int)(r % ntj) * kTJ;
        i0 = (int)(r / ntj) * TI;
    }
};

// worker id with the tiles of one XCD contiguous (blocks are dealt to the 8
// XCDs round robin); gridDim.x is a multiple of 8.
__device__ inline int worker_id() {
    return (int)((blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3));
}

__device__ inline bool inside(int i, int j, int k, int n) {
    return (unsigned)i < (unsigned)n && (unsigned)j < (unsigned)n && (unsigned)k < (unsigned)n;
}

// Stage the (TI+2) x 10 x 34 halo of a tile (channel-last chunks), zeros
// outside, with NT threads.  Each thread issues its loads in batches of up to
// BATCH before the batch's LDS writes (chunks outside the domain load a valid
// address and are replaced by zero), so a tile waits for one memory latency
// per batch, not one per chunk.
template <int TI, int NT, int BATCH>
__device__ inline void stage_halo(const bf16x8* __restrict__ in, int n, int i0, int j0, int k0,
                                  bf16x8* lds) {
    constexpr int NE = (TI + 2) * kHJ * kHK * 3;
    constexpr int PER = (NE + NT - 1) / NT;
    // a rolled loop over batches: unrolled, the tile-invariant index arithmetic
    // of every chunk is hoisted out of the caller's tile loop and held in
    // registers through its compute phase
#pragma unroll 1
    for (int u0 = 0; u0 < PER; u0 += BATCH) {
        bf16x8 v[BATCH];
        bool ok[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            const int e = (int)threadIdx.x + (u0 + u) * NT;
            const int pt = e / 3, c = e - 3 * (e / 3);
            const int hk = pt % kHK, hj = (pt / kHK) % kHJ, hi = pt / (kHK * kHJ);
            const int gi = i0 + hi - 1, gj = j0 + hj - 1, gk = k0 + hk - 1;
            ok[u] = u0 + u < PER && e < NE && inside(gi, gj, gk, n);
            v[u] = in[ok[u] ? (((int64_t)gi * n + gj) * n + gk) * 3 + c : 0];
        }
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            const int e = (int)threadIdx.x + (u0 + u) * NT;
            if (u0 + u < PER && e < NE) lds[e] = ok[u] ? v[u] : bf16x8{};
        }
    }
}

// The halo of a tile in two halves, so that its loads can be in flight while
// the workgroup finishes the previous tile: halo_load issues every chunk load
// of this thread (PER of them, into registers; chunks outside the domain load
// a valid address and are flagged), halo_store writes them to LDS (zeros for
// the flagged ones).
template <int TI, int NT>
struct HaloRegs {
    static constexpr int NE = (TI + 2) * kHJ * kHK * 3;
    static constexpr int PER = (NE + NT - 1) / NT;
    bf16x8 v[PER];
    uint32_t ok = 0;  // bit u: chunk u is inside the domain
    __device__ void load(const bf16x8* __restrict__ in, int n, int i0, int j0, int k0) {
        ok = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            // an empty asm makes the chunk index opaque, so that its (tile
            // invariant) decomposition is recomputed here rather than hoisted
            // out of the tile loop and held in registers through the MFMAs
            int e = (int)threadIdx.x + u * NT;
            asm volatile("" : "+v"(e));
            const int pt = e / 3, c = e - 3 * (e / 3);
            const int hk = pt % kHK, hj = (pt / kHK) % kHJ, hi = pt / (kHK * kHJ);
            const int gi = i0 + hi - 1, gj = j0 + hj - 1, gk = k0 + hk - 1;
            const bool in_d = e < NE && inside(gi, gj, gk, n);
            ok |= in_d ? (1u << u) : 0u;
            v[u] = in[in_d ? (((int64_t)gi * n + gj) * n + gk) * 3 + c : 0];
        }
    }
    __device__ void store(bf16x8* lds) const {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = (int)threadIdx.x + u * NT;
            if (e < NE) lds[e] = (ok >> u) & 1u ? v[u] : bf16x8{};
        }
    }
};

// ---------------------------------------------------------------------------
// 24 -> 24 layer on MFMA.  D[point][co] = sum_k A[point][k] B[k][co] with
// k = (group g = 3*tap + channel group, channel within group), 41 k-steps.
// Workgroup: 8 waves on a tile of 4 x 8 lines (6 x 10 x 34 halo, 98 KB of
// LDS) and the packed weights (B operand, 42 KB of LDS, loaded once per
// persistent workgroup); wave w computes lines 4w .. 4w+3.
//   CE_TRAIN: out0 = gelu(pre), out1 = pre      (pre = acc + bias)
//   CE_INFER: out0 = gelu(pre)
//   CE_BWD:   out0 = acc * gelu'(aux)          (transposed conv of a gradient)
// ---------------------------------------------------------------------------
enum ConvEpi { CE_TRAIN = 0, CE_INFER = 1, CE_BWD = 2 };
constexpr int kCTI = 4;                                   // conv24 tile: 4 x 8 lines
constexpr int kCHalo = (kCTI + 2) * kHJ * kHK;            // 2040 points
constexpr int kConvThreads = 512;
constexpr int kConvWorkers = 256;                         // one workgroup per CU

// After the k-loop the accumulators (+ bias) go through LDS as FP32
// ([wave][4 lines][32 points][24 channels], reusing the halo space) and every
// lane then writes whole 16-byte chunks (8 channels of a point), coalesced.
constexpr int kStgPerWave = 4 * kTK * kC;                         // floats
constexpr int kConvLdsBytes = kCHalo * 3 * 16 > 8 * kStgPerWave * 4 ? kCHalo * 3 * 16
                                                                     : 8 * kStgPerWave * 4;

template <int EPI>
__global__ __launch_bounds__(kConvThreads) void k_conv24(int n, int nsamples,
                                                         const bf16x8* __restrict__ in,
                                                         const bf16x8* __restrict__ wpk,
                                                         const float* __restrict__ bias,
                                                         const bf16x8* __restrict__ aux,
                                                         bf16x8* __restrict__ out0,
                                                         bf16x8* __restrict__ out1) {
    __shared__ __attribute__((aligned(16))) unsigned char smem[kConvLdsBytes];
    __shared__ bf16x8 wl[kSteps * 64];
    bf16x8* lds = reinterpret_cast<bf16x8*>(smem);
    const Tiles<kCTI> T(n, nsamples);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int half = lane >> 5, row = lane & 31;
    float* stg = reinterpret_cast<float*>(smem) + wave * kStgPerWave;
    for (int e = threadIdx.x; e < kSteps * 64; e += kConvThreads) wl[e] = wpk[e];
    const int co = row;
    const float bco = (EPI != CE_BWD && co < kC) ? bias[co] : 0.0f;
    const int64_t N3 = (int64_t)n * n * n;
    // the halo of the next tile is loaded while this one finishes (its
    // epilogue), and written to LDS at the top of the next iteration
    HaloRegs<kCTI, kConvThreads> halo;
    int64_t t = worker_id();
    if (t < T.total) {
        int b, i0, j0, k0;
        T.decode(t, b, i0, j0, k0);
        halo.load(in + (size_t)b * N3 * 3, n, i0, j0, k0);
    }
    for (; t < T.total; t += gridDim.x) {
        int b, i0, j0, k0;
        T.decode(t, b, i0, j0, k0);
        __syncthreads();  // the previous tile's staging reads are done
        halo.store(lds);
        __syncthreads();
        floatx16 acc[4] = {};
        // A operand of step s: chunk cb[q] + group_off(2 s + half) = base[q] +
        // hstep + a compile-time offset (folded into the LDS instruction), so
        // that no per-step address is held in a register
        int base[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int L = wave * 4 + q;
            base[q] = (((L / kTJ + 1) * kHJ + (L % kTJ + 1)) * kHK + row + 1) * 3 + kGroupOffMin;
        }
        const int h1 = half, h94 = half * 94, h808 = half * 808;
        static_for<kSteps>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            constexpr int o = group_off(2 * s) - kGroupOffMin;
            constexpr int d = pair_step(s);
            const int hs = d == 0 ? 0 : d == 1 ? h1 : d == 94 ? h94 : h808;
            const bf16x8 wv = wl[s * 64 + lane];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds[base[q] + hs + o], wv, acc[q],
                                                                 0, 0, 0);
        });
        __syncthreads();  // halo reads done; the space now holds the staging
        if (co < kC) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int pt = 8 * (r >> 2) + 4 * half + (r & 3);
                    stg[(q * kTK + pt) * kC + co] = acc[q][r] + bco;
                }
        }
        __syncthreads();
        // the accumulators are dead: issue the next tile's halo loads now,
        // they complete during this tile's epilogue
        if (t + gridDim.x < T.total) {  // uniform
            int b2, i2, j2, k2;
            T.decode(t + gridDim.x, b2, i2, j2, k2);
            halo.load(in + (size_t)b2 * N3 * 3, n, i2, j2, k2);
        }
        for (int e = lane; e < 4 * kTK * 3; e += 64) {
            const int q = e / (kTK * 3), pt = (e / 3) % kTK, cg = e % 3;
            const int L = wave * 4 + q;
            const int gi = i0 + L / kTJ, gj = j0 + L % kTJ, gk = k0 + pt;
            if (!inside(gi, gj, gk, n)) continue;
            const size_t idx = ((size_t)b * N3 + ((int64_t)gi * n + gj) * n + gk) * 3 + cg;
            const float* v = &stg[(q * kTK + pt) * kC + 8 * cg];
            bf16x8 o0;
            if (EPI == CE_BWD) {
                const bf16x8 pv = aux[idx];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    o0[i] = static_cast<__bf16>(v[i] * gelu_grad_f(static_cast<float>(pv[i])));
            } else {
                bf16x8 o1;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    o1[i] = static_cast<__bf16>(v[i]);
                    o0[i] = static_cast<__bf16>(gelu_f(v[i]));
                }
                if (EPI == CE_TRAIN) out1[idx] = o1;
            }
            out0[idx] = o0;
        }
    }
}

// Pack a mid layer's FP32 weights into the per-lane B operand of k_conv24:
// wpk[s][lane][e] = Wk[co = lane%32][ci = 8*(g%3) + e][tap = g/3], g = 2s + lane/32,
// zero for co >= 24 or g >= 81.  TRANSPOSE packs the weights of the transposed,
// tap-flipped convolution (backward data): Wk[co][ci][tap] = W[ci][co][26 - tap].
template <bool TRANSPOSE>
__global__ __launch_bounds__(64) void k_pack_mid(const float* __restrict__ W, bf16x8* __restrict__ wpk) {
    const int s = blockIdx.x, lane = threadIdx.x;
    const int co = lane & 31, g = 2 * s + (lane >> 5);
    bf16x8 v{};
    if (co < kC && g < kGroups) {
        const int tap = g / 3;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ci = 8 * (g % 3) + e;
            const float w = TRANSPOSE ? W[((int64_t)ci * kC + co) * kTaps + (kTaps - 1 - tap)]
                                      : W[((int64_t)co * kC + ci) * kTaps + tap];
            v[e] = static_cast<__bf16>(w);
        }
    }
    wpk[s * 64 + lane] = v;
}

// ---------------------------------------------------------------------------
// First layer (1 -> 24), VALU FP32: pre[p][co] = b[co] + sum_tap W[co][tap] xn[p + d(tap)],
// xn = x * xs[b].  Thread per point; grid (ceil(N/256), nsamples).
// ---------------------------------------------------------------------------
template <class TIN, bool TRAIN>
__global__ __launch_bounds__(256) void k_conv_first(int n, const TIN* __restrict__ x,
                                                    const float* __restrict__ xs,
                                                    const float* __restrict__ W,
                                                    const float* __restrict__ bias,
                                                    bf16x8* __restrict__ act,
                                                    bf16x8* __restrict__ pre) {
    __shared__ float w[kC * kTaps + kC];
    for (int e = threadIdx.x; e < kC * kTaps; e += 256) w[e] = W[e];
    if (threadIdx.x < kC) w[kC * kTaps + threadIdx.x] = bias[threadIdx.x];
    __syncthreads();
    const int64_t N3 = (int64_t)n * n * n;
    const int b = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= N3) return;
    const int k = (int)(p % n), j = (int)((p / n) % n), i = (int)(p / ((int64_t)n * n));
    const float s = xs[b];
    const TIN* xb = x + (size_t)b * N3;
    // all 27 loads issued before the first use (an index outside the domain
    // loads element 0 and is replaced by zero)
    TIN xr[kTaps];
    bool xok[kTaps];
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) {
        const int ii = i + tap / 9 - 1, jj = j + (tap / 3) % 3 - 1, kk = k + tap % 3 - 1;
        xok[tap] = inside(ii, jj, kk, n);
        xr[tap] = xb[xok[tap] ? ((int64_t)ii * n + jj) * n + kk : 0];
    }
    float xn[kTaps];
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap)
        xn[tap] = xok[tap] ? static_cast<float>(xr[tap]) * s : 0.0f;
    bf16x8* ab = act + ((size_t)b * N3 + p) * 3;
    bf16x8* pb = pre + ((size_t)b * N3 + p) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        bf16x8 av, pv;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int co = 8 * c + e;
            float acc = w[kC * kTaps + co];
#pragma unroll
            for (int tap = 0; tap < kTaps; ++tap) acc = fmaf(w[co * kTaps + tap], xn[tap], acc);
            pv[e] = static_cast<__bf16>(acc);
            av[e] = static_cast<__bf16>(gelu_f(acc));
        }
        ab[c] = av;
        if (TRAIN) pb[c] = pv;
    }
}

// ---------------------------------------------------------------------------
// Last layer (24 -> 1) on the VALU over the LDS halo tile:
// out[b][p] = (b5 + sum_tap sum_ci W[ci][tap] a[p + d(tap)][ci]) * os[b].
// The weights are split once per workgroup into bf16 pairs W = Wh + Wl (about
// 16 bits of each weight) and every activation pair enters two packed bf16
// dot products (v_dot2c_f32_bf16: a.x w.x + a.y w.y + acc), so no activation
// is converted to FP32.  Thread -> points k and k + 16 of a line: consecutive
// lanes read halo points 48 bytes apart (distinct LDS banks) and store to
// consecutive addresses.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

template <class TOUT>
__global__ __launch_bounds__(256) void k_conv_last(int n, int nsamples, const bf16x8* __restrict__ act,
                                                   const float* __restrict__ W,
                                                   const float* __restrict__ bias,
                                                   const float* __restrict__ os,
                                                   TOUT* __restrict__ out) {
    constexpr int kPairs = kC / 2;  // channel pairs per tap
    __shared__ bf16x8 lds[kHalo * 3];
    __shared__ __attribute__((aligned(16))) bf16x2 wh[kTaps * kPairs];  // [tap][pair]
    __shared__ __attribute__((aligned(16))) bf16x2 wl[kTaps * kPairs];
    for (int e = threadIdx.x; e < kTaps * kPairs; e += 256) {
        const int tap = e / kPairs, pr = e % kPairs;
        bf16x2 h, l;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const float w = W[(2 * pr + q) * kTaps + tap];
            const __bf16 hv = static_cast<__bf16>(w);
            h[q] = hv;
            l[q] = static_cast<__bf16>(w - static_cast<float>(hv));
        }
        wh[e] = h;
        wl[e] = l;
    }
    const float b0 = bias[0];
    const Tiles<> T(n, nsamples);
    const int L = threadIdx.x >> 4, kk = threadIdx.x & 15;
    const int64_t N3 = (int64_t)n * n * n;
    for (int64_t t = worker_id(); t < T.total; t += gridDim.x) {
        int b, i0, j0, k0;
        T.decode(t, b, i0, j0, k0);
        __syncthreads();
        stage_halo<kTI, 256, 16>(act + (size_t)b * N3 * 3, n, i0, j0, k0, lds);
        __syncthreads();
        const int base = ((L / kTJ + 1) * kHJ + (L % kTJ + 1)) * kHK + kk + 1;
        float acc0 = b0, acc1 = b0;
#pragma unroll 3
        for (int tap = 0; tap < kTaps; ++tap) {
            const int p0 = (base + tap_off(tap)) * 3;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bf16x8 v0 = lds[p0 + c], v1 = lds[p0 + 48 + c];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int e = tap * kPairs + 4 * c + q;
                    const bf16x2 a0 = {v0[2 * q], v0[2 * q + 1]};
                    const bf16x2 a1 = {v1[2 * q], v1[2 * q + 1]};
                    acc0 = __builtin_amdgcn_fdot2_f32_bf16(a0, wh[e], acc0, false);
                    acc0 = __builtin_amdgcn_fdot2_f32_bf16(a0, wl[e], acc0, false);
                    acc1 = __builtin_amdgcn_fdot2_f32_bf16(a1, wh[e], acc1, false);
                    acc1 = __builtin_amdgcn_fdot2_f32_bf16(a1, wl[e], acc1, false);
                }
            }
        }
        const int gi = i0 + L / kTJ, gj = j0 + L % kTJ, gk = k0 + kk;
        if (gi < n && gj < n) {
            const float sc = os[b];
            const int64_t idx = (int64_t)b * N3 + ((int64_t)gi * n + gj) * n + gk;
            if (gk < n) out[idx] = static_cast<TOUT>(acc0 * sc);
            if (gk + 16 < n) out[idx + 16] = static_cast<TOUT>(acc1 * sc);
        }
    }
}

// Backward data of the last layer: g5[p][ci] = gelu'(pre5[p][ci]) *
// sum_tap W[ci][tap] g6[p - d(tap)].  Thread per point; grid (ceil(N/256), B).
__global__ __launch_bounds__(256) void k_bwd_last(int n, const float* __restrict__ g6,
                                                  const float* __restrict__ W,
                                                  const bf16x8* __restrict__ pre5,
                                                  bf16x8* __restrict__ g5) {
    __shared__ float w[kTaps * kC];  // [tap][ci]
    for (int e = threadIdx.x; e < kTaps * kC; e += 256) {
        const int tap = e / kC, ci = e % kC;
        w[e] = W[ci * kTaps + tap];
    }
    __syncthreads();
    const int64_t N3 = (int64_t)n * n * n;
    const int b = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= N3) return;
    const int k = (int)(p % n), j = (int)((p / n) % n), i = (int)(p / ((int64_t)n * n));
    const float* gb = g6 + (size_t)b * N3;
    float gn[kTaps];
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) {
        const int ii = i - (tap / 9 - 1), jj = j - ((tap / 3) % 3 - 1), kk = k - (tap % 3 - 1);
        gn[tap] = inside(ii, jj, kk, n) ? gb[((int64_t)ii * n + jj) * n + kk] : 0.0f;
    }
    const size_t base = ((size_t)b * N3 + p) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const bf16x8 pv = pre5[base + c];
        bf16x8 gv;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ci = 8 * c + e;
            float acc = 0.0f;
#pragma unroll
            for (int tap = 0; tap < kTaps; ++tap) acc = fmaf(w[tap * kC + ci], gn[tap], acc);
            gv[e] = static_cast<__bf16>(acc * gelu_grad_f(static_cast<float>(pv[e])));
        }
        g5[base + c] = gv;
    }
}

// ---------------------------------------------------------------------------
// Weight gradients.  Every kernel is persistent over the tiles and writes one
// partial per workgroup; k_reduce_partials sums them in a fixed order.
// ---------------------------------------------------------------------------

// Stage the 512 interior points of a tile, channel-last bf16, zeros outside.
__device__ inline void stage_tile(const bf16x8* __restrict__ g, int n, int i0, int j0, int k0,
                                  bf16x8* lds) {
    for (int e = threadIdx.x; e < kTilePts * 3; e += blockDim.x) {
        const int pt = e / 3, c = e - 3 * (e / 3);
        const int L = pt / kTK, kk = pt % kTK;
        const int gi = i0 + L / kTJ, gj = j0 + L % kTJ, gk = k0 + kk;
        bf16x8 v{};
        if (inside(gi, gj, gk, n)) v = g[(((int64_t)gi * n + gj) * n + gk) * 3 + c];
        lds[e] = v;
    }
}

// First layer: dW[co][tap] = sum_p g1[p][co] xn[p + d(tap)], db[co] = sum_p g1[p][co].
// Thread (co = t % 24, group = t / 24 < 10) over the points grp, grp+10, ...
// Partial layout per workgroup: [co][28] (27 taps, bias).
template <class TIN>
__global__ __launch_bounds__(256) void k_wgrad_first(int n, int nsamples, const TIN* __restrict__ x,
                                                     const float* __restrict__ xs,
                                                     const bf16x8* __restrict__ g1,
                                                     float* __restrict__ part) {
    __shared__ float xh[kHalo];
    __shared__ bf16x8 gt[kTilePts * 3];
    __shared__ float red[240 * 28];
    const Tiles<> T(n, nsamples);
    const int t = threadIdx.x, co = t % kC, grp = t / kC;
    float acc[28] = {};
    const int64_t N3 = (int64_t)n * n * n;
    for (int64_t tile = worker_id(); tile < T.total; tile += gridDim.x) {
        int b, i0, j0, k0;
        T.decode(tile, b, i0, j0, k0);
        __syncthreads();
        const float s = xs[b];
        for (int e = t; e < kHalo; e += 256) {
            const int hk = e % kHK, hj = (e / kHK) % kHJ, hi = e / (kHK * kHJ);
            const int gi = i0 + hi - 1, gj = j0 + hj - 1, gk = k0 + hk - 1;
            xh[e] = inside(gi, gj, gk, n)
                        ? static_cast<float>(x[(size_t)b * N3 + ((int64_t)gi * n + gj) * n + gk]) * s
                        : 0.0f;
        }
        stage_tile(g1 + (size_t)b * N3 * 3, n, i0, j0, k0, gt);
        __syncthreads();
        if (grp < 10) {
            const __bf16* gs = reinterpret_cast<const __bf16*>(gt);
            for (int pt = grp; pt < kTilePts; pt += 10) {
                const int L = pt / kTK, kk = pt % kTK;
                const int base = ((L / kTJ + 1) * kHJ + (L % kTJ + 1)) * kHK + kk + 1;
                const float gv = static_cast<float>(gs[pt * kC + co]);
#pragma unroll
                for (int tap = 0; tap < kTaps; ++tap) acc[tap] = fmaf(gv, xh[base + tap_off(tap)], acc[tap]);
                acc[27] += gv;
            }
        }
    }
    __syncthreads();
    if (grp < 10)
        for (int q = 0; q < 28; ++q) red[t * 28 + q] = acc[q];
    __syncthreads();
    for (int e = t; e < kC * 28; e += 256) {
        const int c = e / 28, q = e % 28;
        float s = 0.0f;
        for (int g = 0; g < 10; ++g) s += red[(g * kC + c) * 28 + q];
        part[(size_t)blockIdx.x * kC * 28 + e] = s;
    }
}

// Last layer: dW[ci][tap] = sum_p g6[p] a5[p + d(tap)][ci], db = sum_p g6[p].
// Thread (ci = t % 24, group = t / 24 < 10).  Partial layout [ci][28]; the
// bias sum is kept in ci = 0's slot 27.
__global__ __launch_bounds__(256) void k_wgrad_last(int n, int nsamples, const bf16x8* __restrict__ a5,
                                                    const float* __restrict__ g6,
                                                    float* __restrict__ part) {
    __shared__ bf16x8 lds[kHalo * 3];
    __shared__ float gt[kTilePts];
    __shared__ float red[240 * 28];
    const Tiles<> T(n, nsamples);
    const int t = threadIdx.x, ci = t % kC, grp = t / kC;
    float acc[28] = {};
    const int64_t N3 = (int64_t)n * n * n;
    for (int64_t tile = worker_id(); tile < T.total; tile += gridDim.x) {
        int b, i0, j0, k0;
        T.decode(tile, b, i0, j0, k0);
        __syncthreads();
        stage_halo<kTI, 256, 16>(a5 + (size_t)b * N3 * 3, n, i0, j0, k0, lds);
        for (int e = t; e < kTilePts; e += 256) {
            const int L = e / kTK, kk = e % kTK;
            const int gi = i0 + L / kTJ, gj = j0 + L % kTJ, gk = k0 + kk;
            gt[e] = inside(gi, gj, gk, n) ? g6[(size_t)b * N3 + ((int64_t)gi * n + gj) * n + gk] : 0.0f;
        }
        __syncthreads();
        if (grp < 10) {
            const __bf16* as = reinterpret_cast<const __bf16*>(lds);
            for (int pt = grp; pt < kTilePts; pt += 10) {
                const int L = pt / kTK, kk = pt % kTK;
                const int base = ((L / kTJ + 1) * kHJ + (L % kTJ + 1)) * kHK + kk + 1;
                const float gv = gt[pt];
#pragma unroll
                for (int tap = 0; tap < kTaps; ++tap)
                    acc[tap] = fmaf(gv, static_cast<float>(as[(base + tap_off(tap)) * kC + ci]), acc[tap]);
                acc[27] += gv;
            }
        }
    }
    __syncthreads();
    if (grp < 10)
        for (int q = 0; q < 28; ++q) red[t * 28 + q] = acc[q];
    __syncthreads();
    for (int e = t; e < kC * 28; e += 256) {
        const int c = e / 28, q = e % 28;
        float s = 0.0f;
        for (int g = 0; g < 10; ++g) s += red[(g * kC + c) * 28 + q];
        part[(size_t)blockIdx.x * kC * 28 + e] = s;
    }
}

// Mid layers on MFMA: for dk = blockIdx.y - 1 and the 9 taps (di, dj),
// D_tap[co][ci] = sum_p g[p][co] a[p + (di, dj, dk)][ci] with K = points.
// The g tile and the dk-shifted halo lines of a are staged channel-major in
// LDS (rows padded to 16 mod 256 bytes) so that each lane's 8 consecutive
// points are one aligned 16-byte read.  Wave w takes lines 4w .. 4w+3 of the
// tile (8 k-steps), one accumulator per tap.  The four waves' sums are added
// in LDS in a fixed order; partial layout per workgroup: [3][9][24][24]
// (+ for dk = 0: the bias sums [24] in bpart).
constexpr int kGRow = kTilePts + 8;               // 520 bf16 = 1040 bytes
constexpr int kALines = kHI * kHJ;                // 40 halo lines
constexpr int kARow = kALines * kTK + 8;          // 1288 bf16 = 2576 bytes

__global__ __launch_bounds__(256) void k_wgrad_mid(int n, int nsamples, const bf16x8* __restrict__ g,
                                                   const bf16x8* __restrict__ a,
                                                   float* __restrict__ part,
                                                   float* __restrict__ bpart) {
    __shared__ __attribute__((aligned(16))) __bf16 gs[kC * kGRow];
    __shared__ __attribute__((aligned(16))) __bf16 as[kC * kARow];
    __shared__ float bred[192];
    const int dk = (int)blockIdx.y - 1;
    const Tiles<> T(n, nsamples);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, half = lane >> 5, row = lane & 31;
    floatx16 acc[9] = {};
    float bacc = 0.0f;
    const int64_t N3 = (int64_t)n * n * n;
    const bf16x8 zero{};
    for (int64_t tile = worker_id(); tile < T.total; tile += gridDim.x) {
        int b, i0, j0, k0;
        T.decode(tile, b, i0, j0, k0);
        __syncthreads();
        const bf16x8* gb = g + (size_t)b * N3 * 3;
        const bf16x8* ab = a + (size_t)b * N3 * 3;
        for (int e = t; e < kTilePts * 3; e += 256) {
            const int pt = e / 3, c = e - 3 * (e / 3);
            const int L = pt / kTK, kk = pt % kTK;
            const int gi = i0 + L / kTJ, gj = j0 + L % kTJ, gk = k0 + kk;
            bf16x8 v{};
            if (inside(gi, gj, gk, n)) v = gb[(((int64_t)gi * n + gj) * n + gk) * 3 + c];
#pragma unroll
            for (int q = 0; q < 8; ++q) gs[(8 * c + q) * kGRow + pt] = v[q];
        }
        for (int e = t; e < kALines * kTK * 3; e += 256) {
            const int pt = e / 3, c = e - 3 * (e / 3);
            const int hl = pt / kTK, x = pt % kTK;
            const int gi = i0 + hl / kHJ - 1, gj = j0 + hl % kHJ - 1, gk = k0 + x + dk;
            bf16x8 v{};
            if (inside(gi, gj, gk, n)) v = ab[(((int64_t)gi * n + gj) * n + gk) * 3 + c];
#pragma unroll
            for (int q = 0; q < 8; ++q) as[(8 * c + q) * kARow + pt] = v[q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int L = wave * 4 + q;
            const int li = L / kTJ, lj = L % kTJ;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const int x0 = hh * 16 + 8 * half;
                const bf16x8 av =
                    row < kC ? *reinterpret_cast<const bf16x8*>(&gs[row * kGRow + L * kTK + x0]) : zero;
#pragma unroll
                for (int t9 = 0; t9 < 9; ++t9) {
                    const int hl = (li + t9 / 3) * kHJ + (lj + t9 % 3);
                    const bf16x8 bv =
                        row < kC ? *reinterpret_cast<const bf16x8*>(&as[row * kARow + hl * kTK + x0])
                                 : zero;
                    acc[t9] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t9], 0, 0, 0);
                }
            }
        }
        if (dk == 0 && t < 192) {
            const int c = t >> 3, seg = t & 7;
            float s = 0.0f;
            for (int pt = seg * 64; pt < seg * 64 + 64; ++pt) s += static_cast<float>(gs[c * kGRow + pt]);
            bacc += s;
        }
    }
    // fixed-order sum of the four waves' accumulators through LDS (reuses as[])
    float* fsum = reinterpret_cast<float*>(as);  // 9 x 32 x 32 floats = 36 KB < 61 KB
    for (int w = 0; w < 4; ++w) {
        __syncthreads();
        if (wave == w) {
#pragma unroll
            for (int t9 = 0; t9 < 9; ++t9)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int co = 8 * (r >> 2) + 4 * half + (r & 3);
                    float* dst = &fsum[(t9 * 32 + co) * 32 + row];
                    *dst = (w == 0 ? 0.0f : *dst) + acc[t9][r];
                }
        }
    }
    __syncthreads();
    float* pb = part + ((size_t)blockIdx.x * 3 + (size_t)blockIdx.y) * 9 * kC * kC;
    for (int e = t; e < 9 * kC * kC; e += 256) {
        const int t9 = e / (kC * kC), co = (e / kC) % kC, ci = e % kC;
        pb[e] = fsum[(t9 * 32 + co) * 32 + ci];
    }
    if (dk == 0) {
        if (t < 192) bred[t] = bacc;
        __syncthreads();
        if (t < kC) {
            float s = 0.0f;
            for (int q = 0; q < 8; ++q) s += bred[t * 8 + q];
            bpart[(size_t)blockIdx.x * kC + t] = s;
        }
    }
}

// Sum of per-workgroup partials, scattered into the flat gradient.  One wave
// per output: lane l adds the partials l, l+64, ... in order, then a fixed
// shuffle tree combines the lanes (deterministic).  MODE 0: first layer
// ([co][28] -> W0, b0), MODE 1: last layer ([ci][28] -> W5, b5), MODE 2: mid
// layer ([3][9][24][24] -> W[co][ci][tap]), MODE 3: mid layer bias ([24] -> b).
template <int MODE>
__global__ __launch_bounds__(256) void k_reduce_partials(int G, const float* __restrict__ part,
                                                         float* __restrict__ grad) {
    constexpr int len = MODE == 2 ? 27 * kC * kC : (MODE == 3 ? kC : kC * 28);
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (e >= len) return;  // uniform per wave
    float s = 0.0f;
    for (int w = lane; w < G; w += 64) s += part[(size_t)w * len + e];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane != 0) return;
    if (MODE == 0 || MODE == 1) {
        const int c = e / 28, q = e % 28;
        if (q < 27)
            grad[c * kTaps + q] = s;
        else if (MODE == 0)
            grad[kC * kTaps + c] = s;
        else if (c == 0)
            grad[kC * kTaps] = s;
    } else if (MODE == 2) {
        const int dkI = e / (9 * kC * kC), t9 = (e / (kC * kC)) % 9, co = (e / kC) % kC, ci = e % kC;
        grad[((int64_t)co * kC + ci) * kTaps + t9 * 3 + dkI] = s;
    } else {
        grad[e] = s;
    }
}
template <int MODE>
constexpr int reduce_blocks() {
    return ((MODE == 2 ? 27 * kC * kC : (MODE == 3 ? kC : kC * 28)) + 3) / 4;
}

// ---------------------------------------------------------------------------
// Scales, loss and optimiser.
// ---------------------------------------------------------------------------

// Per-sample partial sums (grid (nblk, B)): MODE 0: sum r^2;
// MODE 1: sum (out - ys[b] z)^2 and sum (ys[b] z)^2 (two partial arrays).
template <class T, int MODE>
__global__ __launch_bounds__(256) void k_sample_sums(int64_t N3, const T* __restrict__ r,
                                                     const float* __restrict__ out,
                                                     const float* __restrict__ ys,
                                                     double* __restrict__ part) {
    const int b = blockIdx.y, nblk = gridDim.x;
    const T* rb = r + (size_t)b * N3;
    double s0 = 0.0, s1 = 0.0;
    if (MODE == 0) {
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N3; i += (int64_t)nblk * 256) {
            const double v = (double)rb[i];
            s0 += v * v;
        }
    } else {
        const float y = ys[b];
        const float* ob = out + (size_t)b * N3;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N3; i += (int64_t)nblk * 256) {
            const double t = (double)((float)rb[i] * y);
            const double d = (double)ob[i] - t;
            s0 += d * d;
            s1 += t * t;
        }
    }
    const double t0 = block_sum<256>(s0);
    __syncthreads();
    const double t1 = MODE == 1 ? block_sum<256>(s1) : 0.0;
    if (threadIdx.x == 0) {
        part[(size_t)b * nblk + blockIdx.x] = t0;
        if (MODE == 1) part[(size_t)(gridDim.y + b) * nblk + blockIdx.x] = t1;
    }
}

// One thread per sample (B <= 64).  MODE 0 (apply): rho = sqrt(sum r^2 / N3);
//   xs = 1/rho, ys = rho / (n+1)^2 (output scale).  MODE 2 (training pair):
//   xs = 1/rho, ys = (n+1)^2 / rho (target scale).  MODE 1 (loss):
//   loss[b] = s0/s1, xs[b] = coef = 2 / (B s1), *mean_out = mean loss.
template <int MODE>
__global__ __launch_bounds__(64) void k_sample_finish(int nblk, int B, int64_t N3, double h2inv,
                                                      const double* __restrict__ part,
                                                      float* __restrict__ xs,
                                                      float* __restrict__ ys,
                                                      double* __restrict__ loss,
                                                      double* __restrict__ mean_out) {
    __shared__ double sl[64];
    const int b = threadIdx.x;
    double l = 0.0;
    if (b < B) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < nblk; ++i) s0 += part[(size_t)b * nblk + i];
        if (MODE == 1) {
            for (int i = 0; i < nblk; ++i) s1 += part[(size_t)(B + b) * nblk + i];
            l = s1 > 0.0 ? s0 / s1 : 0.0;
            loss[b] = l;
            xs[b] = s1 > 0.0 ? (float)(2.0 / (B * s1)) : 0.0f;
        } else {
            const double rho = std::sqrt(s0 / (double)N3);
            xs[b] = rho > 0.0 ? (float)(1.0 / rho) : 0.0f;
            if (MODE == 0)
                ys[b] = (float)(rho / h2inv);
            else
                ys[b] = rho > 0.0 ? (float)(h2inv / rho) : 0.0f;
        }
    }
    if (MODE == 1) {
        sl[b] = l;
        __syncthreads();
        if (b == 0 && mean_out) {
            double s = 0.0;
            for (int q = 0; q < B; ++q) s += sl[q];
            *mean_out = s / B;
        }
    }
}

// Residual loss: e = A7 (out - ys z) with A7 u = 6 u - (sum of the six
// neighbours), zero outside (the grid operator in grid units); partial sums
// of e^2 and (A7 ys z)^2 per sample, as k_sample_sums MODE 1 (grid (nblk, B)).
template <class T>
__global__ __launch_bounds__(256) void k_res_sums(int n, const float* __restrict__ out,
                                                  const T* __restrict__ z,
                                                  const float* __restrict__ ys,
                                                  float* __restrict__ e, double* __restrict__ part) {
    const int b = blockIdx.y, nblk = gridDim.x;
    const int64_t N3 = (int64_t)n * n * n;
    const float y = ys[b];
    const float* ob = out + (size_t)b * N3;
    const T* zb = z + (size_t)b * N3;
    float* eb = e + (size_t)b * N3;
    double s0 = 0.0, s1 = 0.0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < N3; p += (int64_t)nblk * 256) {
        const int k = (int)(p % n), j = (int)((p / n) % n), i = (int)(p / ((int64_t)n * n));
        const float tc = y * static_cast<float>(zb[p]);
        const float dc = ob[p] - tc;
        float dsum = 0.0f, tsum = 0.0f;
        const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                              {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            if (!inside(nb[q][0], nb[q][1], nb[q][2], n)) continue;
            const int64_t o = ((int64_t)nb[q][0] * n + nb[q][1]) * n + nb[q][2];
            const float tz = y * static_cast<float>(zb[o]);
            dsum += ob[o] - tz;
            tsum += tz;
        }
        const float ev = 6.0f * dc - dsum, tv = 6.0f * tc - tsum;
        eb[p] = ev;
        s0 += (double)ev * ev;
        s1 += (double)tv * tv;
    }
    const double t0 = block_sum<256>(s0);
    __syncthreads();
    const double t1 = block_sum<256>(s1);
    if (threadIdx.x == 0) {
        part[(size_t)b * nblk + blockIdx.x] = t0;
        part[(size_t)(gridDim.y + b) * nblk + blockIdx.x] = t1;
    }
}

// g6[b][p] = coef[b] (A7 e)[p]  (A7 is symmetric)
__global__ __launch_bounds__(256) void k_res_grad(int n, const float* __restrict__ e,
                                                  const float* __restrict__ coef,
                                                  float* __restrict__ g6) {
    const int64_t N3 = (int64_t)n * n * n;
    const int b = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= N3) return;
    const int k = (int)(p % n), j = (int)((p / n) % n), i = (int)(p / ((int64_t)n * n));
    const float* eb = e + (size_t)b * N3;
    float s = 6.0f * eb[p];
    if (i > 0) s -= eb[p - (int64_t)n * n];
    if (i + 1 < n) s -= eb[p + (int64_t)n * n];
    if (j > 0) s -= eb[p - n];
    if (j + 1 < n) s -= eb[p + n];
    if (k > 0) s -= eb[p - 1];
    if (k + 1 < n) s -= eb[p + 1];
    g6[(size_t)b * N3 + p] = coef[b] * s;
}

// g6[b][p] = coef[b] (out - ys[b] z)
template <class T>
__global__ __launch_bounds__(256) void k_out_grad(int64_t N3, const float* __restrict__ out,
                                                  const T* __restrict__ z, const float* __restrict__ ys,
                                                  const float* __restrict__ coef, float* __restrict__ g6) {
    const int b = blockIdx.y;
    const float y = ys[b], cf = coef[b];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N3; i += (int64_t)gridDim.x * 256) {
        const size_t o = (size_t)b * N3 + i;
        g6[o] = cf * (out[o] - (float)z[o] * y);
    }
}

// Adam (bias-corrected), FP32 master weights.
__global__ __launch_bounds__(256) void k_adam(int64_t count, float* __restrict__ w, float* __restrict__ m,
                                              float* __restrict__ v, const float* __restrict__ g,
                                              float lr, float b1, float b2, float eps, float bc1,
                                              float bc2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

}  // namespace cnn
}  // namespace numa_hip
