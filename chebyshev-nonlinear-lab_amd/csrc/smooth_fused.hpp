// smooth_fused.hpp - temporally blocked polynomial smoothing: STEPS
// consecutive Jacobi-type / Chebyshev steps
//     d <- c1[s] d + c2[s] (b - A x),   x <- x + d        (s = 0 .. STEPS-1)
// in one kernel that reads x and b once and writes the final x once, instead
// of STEPS kernels that each read x, b, d and write x, d.
//
// Arithmetic is the per-step kernel's (EpPoly in mg.hpp): SmoothArith<T> (the
// storage precision) in the same operation order, x and d rounded to T after
// every step, so the result is bit for bit the one of STEPS separate launches.
//
// Work decomposition.  A workgroup of 1024 threads (16 waves) owns a region
// of 64 (k, one lane each) x 16R (j, R rows per thread) columns and a chunk of
// planes along i.  Few rows per thread keep the register rings small (and out
// of scratch).  It marches over the planes: at plane q it loads x and b,
// then step s advances the column to x_{s+1} at plane q-1-s.  The i-1 / i / i+1
// values of every level live in registers, the j and k neighbours of the plane
// being advanced in a 2-plane LDS ring per level (planes padded with a zero
// border, so every neighbour read is unconditional); one barrier per plane.
// A step is wrong within one cell of the region edge (missing neighbours), so
// the outputs are the columns at least STEPS cells inside the region (or at
// the domain boundary, where the zero Dirichlet values are exact) and the
// planes of the chunk; the halo planes and columns are recomputed by the
// neighbouring workgroups.  Out-of-domain columns and planes hold zeros.  On a
// grid no wider (taller) than the region, one region at the origin spans that
// axis with no halo: both its edges are the boundary, and every node is an
// output (FusedGeom::whole_k / whole_j).
#pragma once

#include <type_traits>
#include <utility>

#include "common.hpp"
#include "stencil.hpp"

namespace numa_hip {

// Arithmetic of the polynomial smoothing steps: the storage precision.
template <class T>
using SmoothArith = std::conditional_t<sizeof(T) == 4, float, double>;

// Arithmetic of one polynomial smoothing step at a node (every smoothing
// kernel uses these, so fused and per-step launches agree bit for bit):
//   s   = fma(6, c, -im) - ip - jm - jp - km - kp      (6c - neighbours)
//   res = fma(-h2inv, s, b)                           (b - A x)
//   d   = c2 res  (first step of a sweep)  |  fma(c1, d, c2 res)
//   x   = c + d
// The smoothers have no counterpart in the reference, so their rounding is
// free to use fused multiply-adds (3 of 13 operations fewer per update).
template <class A>
__device__ __forceinline__ A fma_a(A a, A b, A c) {
    if constexpr (sizeof(A) == 4)
        return __builtin_fmaf(a, b, c);
    else
        return __builtin_fma(a, b, c);
}

template <class A>
__device__ __forceinline__ A poly_res(A b, A h2inv, A c, A im, A ip, A jm, A jp, A km, A kp) {
#pragma clang fp contract(off)
    A s = fma_a<A>(A(6), c, -im);
    s -= ip;
    s -= jm;
    s -= jp;
    s -= km;
    s -= kp;
    return fma_a<A>(-h2inv, s, b);
}

template <class A>
__device__ __forceinline__ A poly_dir(A c1, A d, A c2, A res) {
#pragma clang fp contract(off)
    return fma_a<A>(c1, d, c2 * res);
}

constexpr int kFusedMaxSteps = 6;
constexpr int kFusedThreads = 1024;

struct PolySteps {
    double c1[kFusedMaxSteps];
    double c2[kFusedMaxSteps];
};

template <int STEPS, int R>
struct FusedGeom {
    static constexpr int kW = 64, kH = (kFusedThreads / 64) * R;
    static constexpr int kOutW = kW - 2 * STEPS, kOutH = kH - 2 * STEPS;
    static_assert(kOutW > 0 && kOutH > 0, "halo larger than the region");
    // A region as wide (tall) as the domain needs no halo columns (rows):
    // both of its edges are the zero Dirichlet boundary, so one region at the
    // domain's origin covers that axis and every node in it is an output.
    static constexpr bool whole_k(int n) { return n <= kW; }
    static constexpr bool whole_j(int n) { return n <= kH; }
    static constexpr int tiles_k(int n) { return whole_k(n) ? 1 : (n + kOutW - 1) / kOutW; }
    static constexpr int tiles_j(int n) { return whole_j(n) ? 1 : (n + kOutH - 1) / kOutH; }
};

// Planes of the b register ring and of the unrolled march (see k_poly_fused):
// STEPS + 1 planes where lcm(6, STEPS + 1) <= 12, else 6 or 12.
constexpr int fused_b_ring(int steps) {
    return steps <= 3 ? steps + 1 : steps <= 5 ? 6 : 12;
}
constexpr int fused_unroll(int steps) { return steps == 3 || steps == 6 ? 12 : 6; }

// Raw buffer access (MUBUF) to one plane: a uniform descriptor (base, size in
// bytes) and a 32-bit per-thread byte offset.  Loads past the size return
// zero and stores past it are dropped, so out-of-domain rows and planes need
// no branches: their offset (kNoOffset) or size (0) is out of range.
constexpr uint32_t kNoOffset = 0x80000000u;
constexpr int kBufWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, kBufWord3);
}

template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

template <class T>
__device__ __forceinline__ void buf_store(T v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(__attribute__((ext_vector_type(2))) uint32_t, v), r, (int)off, 0, 0);
}

// The value of the lane one to the left (wave_shr:1) or right (wave_shl:1)
// in the wave, 0 for lane 0 / lane 63: a DPP move, no LDS access.  The wave
// spans the region's 64 k columns, so these are the k - 1 / k + 1 neighbours,
// and the region edge reads 0 as the LDS ring's zero border does.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true);
}
template <int CTRL, class A>
__device__ __forceinline__ A lane_shift(A v) {
    if constexpr (sizeof(A) == 4) {
        return __builtin_bit_cast(A, dpp_u32<CTRL>(__builtin_bit_cast(uint32_t, v)));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint64_t lo = dpp_u32<CTRL>((uint32_t)u), hi = dpp_u32<CTRL>((uint32_t)(u >> 32));
        return __builtin_bit_cast(A, lo | (hi << 32));
    }
}
constexpr int kDppWaveShr1 = 0x138;  // lane l reads lane l - 1
constexpr int kDppWaveShl1 = 0x130;  // lane l reads lane l + 1

// Waves per SIMD the register allocation must allow: 8 (two 16-wave
// workgroups per CU, at most 64 VGPRs) for the variants whose rings fit that
// without spilling, else no bound.
template <class T, int STEPS, int R, bool NORM>
constexpr int fused_waves_per_eu() {
    constexpr int words = (int)(sizeof(T) / 4) * R;  // 32-bit registers per ring value
    if (NORM) return 1;  // the FP64 accumulator: no bound
    return (words <= 2 && STEPS <= 2) || (words <= 4 && STEPS == 1) ? 8 : 1;
}

// NORM: every workgroup also writes the sum of squares (FP64) of the x it
// stores to norm_part[blockIdx.x] (0 for workgroups without a tile).  Used
// with the smallest rows per thread and at most 3 steps (the variants that
// hold the accumulator without spilling).
template <class T, int STEPS, int R, bool NORM = false>
__device__ __forceinline__ void poly_fused_body(GridPos g, int n, int ichunk, int tiles_k,
                                                int tiles_j, int chunks, double h2inv_d,
                                                const PolySteps& cs, const T* __restrict__ x,
                                                const T* __restrict__ b, T* __restrict__ xo,
                                                double* __restrict__ norm_part) {
#pragma clang fp contract(off)
    using A = SmoothArith<T>;
    using G = FusedGeom<STEPS, R>;
    constexpr int W = G::kW, H = G::kH;
    // Register rings are indexed by the plane's position t = q - q_first in
    // the march: x of each level by t mod 3, d by t mod 2, b by t mod BR
    // (BR >= STEPS + 1 planes).  The march is unrolled by U planes, a
    // multiple of all three periods, so every ring slot is a compile-time
    // register and no value moves between planes.
    constexpr int BR = fused_b_ring(STEPS);
    constexpr int U = fused_unroll(STEPS);
    static_assert(U % 6 == 0 && U % BR == 0 && BR > STEPS, "ring periods");
    // LDS ring: level s (s < STEPS), 2 planes (by t mod 2) of (H+2) x (W+2)
    // with a zero border
    __shared__ T ring[STEPS][2][H + 2][W + 2];

    // XCD-aware block -> tile: the blocks an XCD receives (every 8th) walk
    // consecutive tiles, k fastest, so overlapping halos meet in one L2.
    const int G8 = (int)(g.nb >> 3);
    const int lin = (int)(g.bid & 7) * G8 + (int)(g.bid >> 3);
    const int total = tiles_k * tiles_j * chunks;
    if (lin >= total) {  // uniform per block, before any barrier
        if (NORM && threadIdx.x == 0) norm_part[g.bid] = 0.0;
        return;
    }
    {
        T* flat = &ring[0][0][0][0];
        for (int e = threadIdx.x; e < STEPS * 2 * (H + 2) * (W + 2); e += kFusedThreads)
            flat[e] = T(0);
    }
    const int tk = lin % tiles_k, tj = (lin / tiles_k) % tiles_j, ch = lin / (tiles_k * tiles_j);
    const bool wk = G::whole_k(n), wj = G::whole_j(n);
    const int k0 = wk ? 0 : tk * G::kOutW - STEPS, j0 = wj ? 0 : tj * G::kOutH - STEPS;
    const int ib = ch * ichunk, ie = min(n, ib + ichunk);

    const A h2inv = static_cast<A>(h2inv_d);
    A c1[STEPS], c2[STEPS];
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
        c1[s] = static_cast<A>(cs.c1[s]);
        c2[s] = static_cast<A>(cs.c2[s]);
    }

    const int lane = threadIdx.x & 63, ty = threadIdx.x >> 6;
    A h_row[R];  // h2inv in the domain, 0 for columns outside it
    const int k = k0 + lane;
    const int64_t nn = (int64_t)n * n;  // plane stride
    bool col_in[R];
    int jrow[R];
    // byte offset of the row's column within a plane (n <= 2048: < 2^25), for
    // the loads (kNoOffset outside the domain) and for the output store
    // (kNoOffset unless an output column of this region)
    const uint32_t plane_bytes = (uint32_t)nn * (uint32_t)sizeof(T);
    const bool out_k = wk || (lane >= STEPS && lane < W - STEPS);
    uint32_t off[R], off_out[R];
    T out_mask[R];  // NORM: 1 for this thread's output columns, else 0
#pragma unroll
    for (int r = 0; r < R; ++r) {
        jrow[r] = ty + (kFusedThreads / 64) * r;
        const int j = j0 + jrow[r];
        col_in[r] = (unsigned)j < (unsigned)n && (unsigned)k < (unsigned)n;
        h_row[r] = col_in[r] ? h2inv : A(0);
        off[r] = col_in[r] ? (uint32_t)(j * n + k) * (uint32_t)sizeof(T) : kNoOffset;
        off_out[r] = out_k && (wj || (jrow[r] >= STEPS && jrow[r] < H - STEPS)) ? off[r]
                                                                                  : kNoOffset;
        out_mask[r] = off_out[r] != kNoOffset ? T(1) : T(0);
    }

    T xs[STEPS][R][3];
    T ds[STEPS][R][2];
    T bb[BR][R];
#pragma unroll
    for (int s = 0; s < STEPS; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            xs[s][r][0] = xs[s][r][1] = xs[s][r][2] = T(0);
            ds[s][r][0] = ds[s][r][1] = T(0);
        }
#pragma unroll
    for (int s = 0; s < BR; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r) bb[s][r] = T(0);

    // plane q of x (zero if x is null) and b; zeros outside the domain
    auto load_plane = [&](int q, T (&vx)[R], T (&vb)[R]) {
        const bool q_in = (unsigned)q < (unsigned)n;
        const int64_t base = (int64_t)(q_in ? q : 0) * nn;
        const uint32_t bytes = q_in ? plane_bytes : 0u;
        const __amdgpu_buffer_rsrc_t rb = plane_rsrc(b + base, bytes);
        const __amdgpu_buffer_rsrc_t rx = plane_rsrc(x ? x + base : b, x ? bytes : 0u);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            vx[r] = buf_load<T>(rx, off[r]);
            vb[r] = buf_load<T>(rb, off[r]);
        }
    };

    const int q_first = ib - STEPS, q_last = ie + STEPS - 1;
    const int nplanes = q_last - q_first + 1;
    double nsq = 0.0;  // NORM: squares of the stored x
    T nx[R], nb[R];
    load_plane(q_first, nx, nb);
    __syncthreads();  // zero borders written

    // One plane of the march, t = q - q_first, PH = t mod U.
    auto plane = [&](auto ph, int q) {
        constexpr int PH = decltype(ph)::value;
        // The j neighbours of every step's plane were written in the previous
        // iteration (before the barrier that ended it), and this iteration
        // writes the other ring slots: all reads are issued first, so their
        // latency overlaps the writes, the loads and the steps' VALU work.
        T jmv[STEPS][R], jpv[STEPS][R];
        static_for<STEPS>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            constexpr int tp = PH - 1 - s + 6 * U;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                jmv[s][r] = ring[s][tp & 1][jrow[r]][lane + 1];
                jpv[s][r] = ring[s][tp & 1][jrow[r] + 2][lane + 1];
            }
        });
        // take in the loaded plane q
#pragma unroll
        for (int r = 0; r < R; ++r) {
            bb[PH % BR][r] = nb[r];
            xs[0][r][PH % 3] = nx[r];
            ring[0][PH & 1][jrow[r] + 1][lane + 1] = nx[r];
        }
        if (q < q_last) load_plane(q + 1, nx, nb);  // in flight during the steps
        static_for<STEPS>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            constexpr int tp = PH - 1 - s + 6 * U;  // t of the plane advanced (mod U), >= 0
            const int p = q - 1 - s;
            const bool p_in = (unsigned)p < (unsigned)n;  // uniform
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int jr = jrow[r];
                const A c = static_cast<A>(xs[s][r][tp % 3]);
                const A im = static_cast<A>(xs[s][r][(tp - 1) % 3]);
                const A ip = static_cast<A>(xs[s][r][(tp + 1) % 3]);
                const A jm = static_cast<A>(jmv[s][r]);
                const A jp = static_cast<A>(jpv[s][r]);
                const A km = lane_shift<kDppWaveShr1>(c);
                const A kp = lane_shift<kDppWaveShl1>(c);
                // h = 0 outside the domain (columns: h_row; planes: p_in), where
                // b = 0 and x = 0: then res = 0, d = 0 and x stays exactly 0
                const A h = p_in ? h_row[r] : A(0);
                const A res =
                    poly_res<A>(static_cast<A>(bb[tp % BR][r]), h, c, im, ip, jm, jp, km, kp);
                A dn;
                if constexpr (s == 0)
                    dn = c2[0] * res;
                else
                    dn = poly_dir<A>(c1[s], static_cast<A>(ds[s - 1][r][tp & 1]), c2[s], res);
                const T xn = static_cast<T>(c + dn);
                if constexpr (s + 1 < STEPS) {
                    ds[s][r][tp & 1] = static_cast<T>(dn);
                    xs[s + 1][r][tp % 3] = xn;
                    ring[s + 1][tp & 1][jr + 1][lane + 1] = xn;
                } else {
                    // output planes [ib, ie) of the chunk, output columns
                    const bool p_out = p >= ib && p < ie;
                    const __amdgpu_buffer_rsrc_t ro =
                        plane_rsrc(xo + (int64_t)(p_out ? p : 0) * nn, p_out ? plane_bytes : 0u);
                    buf_store<T>(xn, ro, off_out[r]);
                    if constexpr (NORM) {
                        if (p_out) {  // uniform; x * 1 and x * 0 are exact
                            const double v = static_cast<double>(xn * out_mask[r]);
                            nsq = __builtin_fma(v, v, nsq);
                        }
                    }
                }
            }
        });
        __syncthreads();
    };
    // whole groups of U planes without tests (no value moves between
    // registers), then the remaining planes
    int t0 = 0;
    for (; t0 + U <= nplanes; t0 += U)
        static_for<U>([&](auto ph) { plane(ph, q_first + t0 + decltype(ph)::value); });
    static_for<U>([&](auto ph) {
        constexpr int PH = decltype(ph)::value;
        if (t0 + PH < nplanes) plane(ph, q_first + t0 + PH);
    });
    if constexpr (NORM) {
        const double t = block_sum<kFusedThreads>(nsq);
        if (threadIdx.x == 0) norm_part[g.bid] = t;
    }
}
template <class T, int STEPS, int R, bool NORM = false>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kFusedThreads),
                          amdgpu_waves_per_eu(fused_waves_per_eu<T, STEPS, R, NORM>()))) void
k_poly_fused(int n, int ichunk, int tiles_k, int tiles_j, int chunks, double h2inv_d, PolySteps cs,
             const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ xo,
             double* __restrict__ norm_part) {
    poly_fused_body<T, STEPS, R, NORM>(hw_grid(), n, ichunk, tiles_k, tiles_j, chunks, h2inv_d, cs,
                                       x, b, xo, norm_part);
}

// Launch geometry of k_poly_fused for an n^3 level.
template <int STEPS, int R>
inline dim3 fused_grid(int n, int ichunk, int& tiles_k, int& tiles_j, int& chunks) {
    using G = FusedGeom<STEPS, R>;
    tiles_k = G::tiles_k(n);
    tiles_j = G::tiles_j(n);
    chunks = (n + ichunk - 1) / ichunk;
    const int64_t total = (int64_t)tiles_k * tiles_j * chunks;
    return dim3((unsigned)((total + 7) / 8 * 8));
}

// ---------------------------------------------------------------------------
// Box form of the same STEPS smoothing steps, for the small and middle levels
// of the hierarchy, where the march above is latency-bound (few workgroups,
// one barrier per plane along a chain of chunk + 2 STEPS planes).
//
// A workgroup of 512 threads owns a box of EI x EJ x EK output nodes.  It
// loads the box grown by STEPS nodes on every side (x, b; zeros outside the
// domain) at once, so all its loads are in flight together, keeps x in a
// double-buffered LDS copy of the grown box and b, d, x of its own nodes in
// registers (node e = threadIdx.x + 512 m, box order k fastest), and runs the
// steps with one barrier each.  Step s advances the nodes at least s + 1 from
// the box surface; their neighbours are at least s from it, which step s - 1
// advanced (or the load, for s = 0), so every LDS read is unconditional and
// in the box.  The last step advances exactly the output nodes; they go
// through LDS to coalesced stores.  Arithmetic and roundings are those of
// k_poly_fused.
constexpr int kBoxThreads = 512;

template <int STEPS, int EI, int EJ, int EK>
struct BoxGeom {
    static constexpr int BI = EI + 2 * STEPS, BJ = EJ + 2 * STEPS, BK = EK + 2 * STEPS;
    static constexpr int kNodes = BI * BJ * BK;
    static constexpr int kPer = (kNodes + kBoxThreads - 1) / kBoxThreads;
};

template <class T, int STEPS, int EI, int EJ, int EK, bool NORM = false>
__device__ __forceinline__ void poly_box_body(GridPos g, int n, int tiles_k, int tiles_j,
                                              int total, double h2inv_d, const PolySteps& cs,
                                              const T* __restrict__ x, const T* __restrict__ b,
                                              T* __restrict__ xo, double* __restrict__ norm_part) {
#pragma clang fp contract(off)
    using A = SmoothArith<T>;
    using G = BoxGeom<STEPS, EI, EJ, EK>;
    constexpr int BI = G::BI, BJ = G::BJ, BK = G::BK, NB = G::kNodes, P = G::kPer;
    constexpr int SI = BJ * BK;  // i stride of the box
    __shared__ T buf[2][NB];

    // XCD-aware block -> box (k fastest), as in k_poly_fused
    const int G8 = (int)(g.nb >> 3);
    const int lin = (int)(g.bid & 7) * G8 + (int)(g.bid >> 3);
    if (lin >= total) {  // uniform per block, before any barrier
        if (NORM && threadIdx.x == 0) norm_part[g.bid] = 0.0;
        return;
    }
    const int tk = lin % tiles_k, tj = (lin / tiles_k) % tiles_j, ti = lin / (tiles_k * tiles_j);
    const int i0 = ti * EI - STEPS, j0 = tj * EJ - STEPS, k0 = tk * EK - STEPS;

    const A h2inv = static_cast<A>(h2inv_d);
    A c1[STEPS], c2[STEPS];
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
        c1[s] = static_cast<A>(cs.c1[s]);
        c2[s] = static_cast<A>(cs.c2[s]);
    }
    // the box's planes in the domain, [ia, iz): one buffer descriptor each
    // for x and b (a span of at most BI planes: < 2^32 bytes for n <= 2048)
    const int64_t nn = (int64_t)n * n;
    const uint32_t row_bytes = (uint32_t)n * (uint32_t)sizeof(T);
    const uint32_t plane_bytes = (uint32_t)nn * (uint32_t)sizeof(T);
    const int ia = max(i0, 0), iz = min(i0 + BI, n);
    const uint32_t span = (uint32_t)(iz - ia) * plane_bytes;
    const __amdgpu_buffer_rsrc_t rb = plane_rsrc(b + ia * nn, span);
    const __amdgpu_buffer_rsrc_t rx = plane_rsrc(x ? x + ia * nn : b, x ? span : 0u);

    // per owned node: x, b, d and its depth (distance to the box surface;
    // 0 for slots past the box, which no step advances) | 256 if in the domain
    T xv[P], bv[P], dv[P];
    int info[P];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int e = (int)threadIdx.x + m * kBoxThreads;
        const int ii = e / SI, jj = (e / BK) % BJ, kk = e % BK;
        const int gi = i0 + ii, gj = j0 + jj, gk = k0 + kk;
        const bool node = e < NB;
        const bool in = node && (unsigned)gi < (unsigned)n && (unsigned)gj < (unsigned)n &&
                        (unsigned)gk < (unsigned)n;
        const int dep = min(min(min(ii, BI - 1 - ii), min(jj, BJ - 1 - jj)), min(kk, BK - 1 - kk));
        info[m] = node ? (dep | (in ? 256 : 0)) : 0;
        const uint32_t off = in ? (uint32_t)(gi - ia) * plane_bytes + (uint32_t)gj * row_bytes +
                                      (uint32_t)gk * (uint32_t)sizeof(T)
                                : kNoOffset;
        xv[m] = buf_load<T>(rx, off);
        bv[m] = buf_load<T>(rb, off);
        dv[m] = T(0);
    }
    // LDS writes after every load is in flight (a write next to its load made
    // each slot wait for its own load)
#pragma unroll
    for (int m = 0; m < P; ++m) {
        const int e = (int)threadIdx.x + m * kBoxThreads;
        if (e < NB) buf[0][e] = xv[m];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
        const T* cur = buf[s & 1];
        T* nxt = buf[(s + 1) & 1];
#pragma unroll
        for (int m = 0; m < P; ++m) {
            if ((info[m] & 255) < s + 1) continue;
            const int e = (int)threadIdx.x + m * kBoxThreads;
            const A c = static_cast<A>(xv[m]);
            const A im = static_cast<A>(cur[e - SI]), ip = static_cast<A>(cur[e + SI]);
            const A jm = static_cast<A>(cur[e - BK]), jp = static_cast<A>(cur[e + BK]);
            const A km = static_cast<A>(cur[e - 1]), kp = static_cast<A>(cur[e + 1]);
            const A res = poly_res<A>(static_cast<A>(bv[m]), h2inv, c, im, ip, jm, jp, km, kp);
            const A dn = s == 0 ? c2[0] * res : poly_dir<A>(c1[s], static_cast<A>(dv[m]), c2[s], res);
            const bool live = (info[m] & 256) != 0;
            const T xn = live ? static_cast<T>(c + dn) : T(0);
            dv[m] = live ? static_cast<T>(dn) : T(0);
            xv[m] = xn;
            nxt[e] = xn;
        }
        __syncthreads();
    }
    // output nodes (the box shrunk by STEPS), k fastest, from the last LDS copy
    const T* fin = buf[STEPS & 1];
    const int oi = ti * EI, oj = tj * EJ, ok = tk * EK;
    const int oz = min(oi + EI, n);
    const __amdgpu_buffer_rsrc_t ro = plane_rsrc(xo + oi * nn, (uint32_t)(oz - oi) * plane_bytes);
    constexpr int NO = EI * EJ * EK;
    double nsq = 0.0;  // NORM: squares of the stored x
#pragma unroll
    for (int o = (int)threadIdx.x; o < NO; o += kBoxThreads) {
        const int ii = o / (EJ * EK), jj = (o / EK) % EJ, kk = o % EK;
        const bool in = oj + jj < n && ok + kk < n;  // planes past n: out of the span
        const uint32_t off = in ? (uint32_t)ii * plane_bytes + (uint32_t)(oj + jj) * row_bytes +
                                      (uint32_t)(ok + kk) * (uint32_t)sizeof(T)
                                : kNoOffset;
        const T v = fin[((ii + STEPS) * BJ + jj + STEPS) * BK + kk + STEPS];
        buf_store<T>(v, ro, off);
        if constexpr (NORM) {
            const double d = in && oi + ii < n ? (double)v : 0.0;
            nsq += d * d;
        }
    }
    if constexpr (NORM) {
        const double t = block_sum<kBoxThreads>(nsq);
        if (threadIdx.x == 0) norm_part[g.bid] = t;
    }
}
template <class T, int STEPS, int EI, int EJ, int EK, bool NORM = false>
__global__ __launch_bounds__(kBoxThreads) void k_poly_box(int n, int tiles_k, int tiles_j,
                                                          int total, double h2inv_d, PolySteps cs,
                                                          const T* __restrict__ x,
                                                          const T* __restrict__ b,
                                                          T* __restrict__ xo,
                                                          double* __restrict__ norm_part) {
    poly_box_body<T, STEPS, EI, EJ, EK, NORM>(hw_grid(), n, tiles_k, tiles_j, total, h2inv_d, cs, x,
                                              b, xo, norm_part);
}

// Launch geometry of k_poly_box for an n^3 level.
template <int EI, int EJ, int EK>
inline dim3 box_grid(int n, int& tiles_k, int& tiles_j, int& total) {
    tiles_k = (n + EK - 1) / EK;
    tiles_j = (n + EJ - 1) / EJ;
    const int tiles_i = (n + EI - 1) / EI;
    total = tiles_k * tiles_j * tiles_i;
    return dim3((unsigned)((total + 7) / 8 * 8));
}

}  // namespace numa_hip
