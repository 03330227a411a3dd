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
// planes along i.  Few rows per thread keep the register rings small, so two
// such workgroups (32 waves) fit a CU and keep enough loads in flight.  It marches over the planes: at plane q it loads x and b,
// then step s advances the column to x_{s+1} at plane q-1-s.  The i-1 / i / i+1
// values of every level live in registers, the j and k neighbours of the plane
// being advanced in a 2-plane LDS ring per level (planes padded with a zero
// border, so every neighbour read is unconditional); one barrier per plane.
// A step is wrong within one cell of the region edge (missing neighbours), so
// the outputs are the columns at least STEPS cells inside the region (or at
// the domain boundary, where the zero Dirichlet values are exact) and the
// planes of the chunk; the halo planes and columns are recomputed by the
// neighbouring workgroups.  Out-of-domain columns and planes hold zeros.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "stencil.hpp"

namespace numa_hip {

// Arithmetic of the polynomial smoothing steps: the storage precision.
template <class T>
using SmoothArith = std::conditional_t<sizeof(T) == 4, float, double>;

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
};

template <class T, int STEPS, int R>
__global__ __launch_bounds__(kFusedThreads) void k_poly_fused(int n, int ichunk, int tiles_k,
                                                              int tiles_j, int chunks,
                                                              double h2inv_d, PolySteps cs,
                                                              const T* __restrict__ x,
                                                              const T* __restrict__ b,
                                                              T* __restrict__ xo) {
#pragma clang fp contract(off)
    using A = SmoothArith<T>;
    using G = FusedGeom<STEPS, R>;
    constexpr int W = G::kW, H = G::kH;
    // LDS ring: level s (s < STEPS), 2 planes of (H+2) x (W+2) with a zero border
    __shared__ T ring[STEPS][2][H + 2][W + 2];

    // XCD-aware block -> tile: the blocks an XCD receives (every 8th) walk
    // consecutive tiles, k fastest, so overlapping halos meet in one L2.
    const int G8 = gridDim.x >> 3;
    const int lin = (int)(blockIdx.x & 7) * G8 + (int)(blockIdx.x >> 3);
    const int total = tiles_k * tiles_j * chunks;
    if (lin >= total) return;  // uniform per block, before any barrier
    {
        T* flat = &ring[0][0][0][0];
        for (int e = threadIdx.x; e < STEPS * 2 * (H + 2) * (W + 2); e += kFusedThreads)
            flat[e] = T(0);
    }
    const int tk = lin % tiles_k, tj = (lin / tiles_k) % tiles_j, ch = lin / (tiles_k * tiles_j);
    const int k0 = tk * G::kOutW - STEPS, j0 = tj * G::kOutH - STEPS;
    const int ib = ch * ichunk, ie = min(n, ib + ichunk);

    const A h2inv = static_cast<A>(h2inv_d);
    A c1[STEPS], c2[STEPS];
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
        c1[s] = static_cast<A>(cs.c1[s]);
        c2[s] = static_cast<A>(cs.c2[s]);
    }

    const int lane = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int k = k0 + lane;
    const int64_t N1 = n;
    bool col_in[R];
    int jrow[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        jrow[r] = ty + (kFusedThreads / 64) * r;
        const int j = j0 + jrow[r];
        col_in[r] = (unsigned)j < (unsigned)n && (unsigned)k < (unsigned)n;
    }
    const bool out_k = lane >= STEPS && lane < W - STEPS;

    // register rings (per row): xs[s][0..2] = x_s at planes (q-s-2, q-s-1, q-s),
    // ds[s][0..1] = d produced by step s at planes (q-2-s, q-1-s),
    // bb[0..STEPS] = b at planes (q, q-1, ..., q-STEPS).
    T xs[STEPS][R][3];
    T ds[STEPS][R][2];
    T bb[STEPS + 1][R];
#pragma unroll
    for (int s = 0; s < STEPS; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            xs[s][r][0] = xs[s][r][1] = xs[s][r][2] = T(0);
            ds[s][r][0] = ds[s][r][1] = T(0);
        }
#pragma unroll
    for (int s = 0; s <= STEPS; ++s)
#pragma unroll
        for (int r = 0; r < R; ++r) bb[s][r] = T(0);

    auto load_plane = [&](int q, T (&vx)[R], T (&vb)[R]) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            vx[r] = T(0);
            vb[r] = T(0);
            if (col_in[r] && (unsigned)q < (unsigned)n) {
                const int64_t id = ((int64_t)q * N1 + (j0 + jrow[r])) * N1 + k;
                if (x) vx[r] = x[id];
                vb[r] = b[id];
            }
        }
    };

    const int q_first = ib - STEPS, q_last = ie + STEPS - 1;
    T nx[R], nb[R];
    load_plane(q_first, nx, nb);
    __syncthreads();  // zero borders written
    for (int q = q_first; q <= q_last; ++q) {
        // shift the rings by one plane and take in the loaded plane q
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int s = STEPS; s > 0; --s) bb[s][r] = bb[s - 1][r];
            bb[0][r] = nb[r];
#pragma unroll
            for (int s = 0; s < STEPS; ++s) {
                xs[s][r][0] = xs[s][r][1];
                xs[s][r][1] = xs[s][r][2];
                ds[s][r][0] = ds[s][r][1];
            }
            xs[0][r][2] = nx[r];
            ring[0][q & 1][jrow[r] + 1][lane + 1] = nx[r];
        }
        if (q < q_last) load_plane(q + 1, nx, nb);  // in flight during the steps
#pragma unroll
        for (int s = 0; s < STEPS; ++s) {
            const int p = q - 1 - s;  // plane advanced by step s
            const bool p_in = (unsigned)p < (unsigned)n;
            const int slot = p & 1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int jr = jrow[r];
                const A c = static_cast<A>(xs[s][r][1]);
                const A im = static_cast<A>(xs[s][r][0]), ip = static_cast<A>(xs[s][r][2]);
                const A jm = static_cast<A>(ring[s][slot][jr][lane + 1]);
                const A jp = static_cast<A>(ring[s][slot][jr + 2][lane + 1]);
                const A km = static_cast<A>(ring[s][slot][jr + 1][lane]);
                const A kp = static_cast<A>(ring[s][slot][jr + 1][lane + 2]);
                const A res =
                    static_cast<A>(bb[s + 1][r]) - h2inv * lap_num<A>(c, im, ip, jm, jp, km, kp);
                const A dn = s == 0 ? c2[0] * res
                                    : c1[s] * static_cast<A>(ds[s - 1][r][0]) + c2[s] * res;
                const bool live = p_in && col_in[r];
                const T xn = live ? static_cast<T>(c + dn) : T(0);
                if (s + 1 < STEPS) {
                    ds[s][r][1] = live ? static_cast<T>(dn) : T(0);
                    xs[s + 1][r][2] = xn;
                    ring[s + 1][(q - 1 - s) & 1][jr + 1][lane + 1] = xn;
                } else if (live && out_k && jr >= STEPS && jr < H - STEPS && p >= ib && p < ie) {
                    xo[((int64_t)p * N1 + (j0 + jr)) * N1 + k] = xn;
                }
            }
        }
        __syncthreads();
    }
}

// Launch geometry of k_poly_fused for an n^3 level.
template <int STEPS, int R>
inline dim3 fused_grid(int n, int ichunk, int& tiles_k, int& tiles_j, int& chunks) {
    using G = FusedGeom<STEPS, R>;
    tiles_k = (n + G::kOutW - 1) / G::kOutW;
    tiles_j = (n + G::kOutH - 1) / G::kOutH;
    chunks = (n + ichunk - 1) / ichunk;
    const int64_t total = (int64_t)tiles_k * tiles_j * chunks;
    return dim3((unsigned)((total + 7) / 8 * 8));
}

}  // namespace numa_hip
