// stencil.hpp - 7-point stencil kernels of Bratu3D on gfx950.
//
// Layout: interior nodes only, id = (i*n + j)*n + k, k fastest (the reference's
// layout, problems/bratu3d.hpp:85), zero Dirichlet values outside.
//
// One generic driver k_stencil: a block is 64 (k) x 4 (j) threads (32 x 8 for
// n <= 32).  Either one node per thread, or each thread owns one (j,k) column
// of a slab of `ichunk` planes and marches along i, keeping the i-1 / i / i+1
// values in registers and prefetching the next plane's loads before computing
// the current one.  The j+-1 and
// k+-1 neighbours are re-read and hit the CU's L1 / the XCD's L2 (the
// neighbouring waves and blocks stream the same rows), so HBM traffic stays
// close to one read per input field.  A Loader gives the field value(s) at a
// node; an Epilogue combines centre and neighbours, writes the outputs and
// returns the node's contribution to an optional fused FP64 reduction.
//
// Arithmetic follows the reference's operation order (bratu3d.hpp:80-94,
// operator.hpp:101-134) with FMA contraction off, so the only difference to
// the CPU is the device exp().
#pragma once

#include "common.hpp"

namespace numa_hip {

struct P1 {
    double a;
    __device__ static P1 zero() { return {0.0}; }
};
struct P2 {
    double a, b;
    __device__ static P2 zero() { return {0.0, 0.0}; }
};

template <class T>
__device__ inline double ld(const T* p, int64_t i) {
    return static_cast<double>(p[i]);
}

// -Lap_h numerator in the reference's order: 6c - im - ip - jm - jp - km - kp
// (in the arithmetic type A).
template <class A>
__device__ inline A lap_num(A c, A im, A ip, A jm, A jp, A km, A kp) {
#pragma clang fp contract(off)
    A s = A(6) * c;
    s -= im;
    s -= ip;
    s -= jm;
    s -= jp;
    s -= km;
    s -= kp;
    return s;
}

// Loader:   Raw load(id) issues the loads of one node (no branches),
//           Val make(raw) combines them into the field value(s).
// Epilogue: Extra preload(id) issues the node's other loads (F0, b, d, ...),
//           operator()(id, extra, centre, 6 neighbours) computes, stores and
//           returns the node's reduction contribution.
// Out-of-domain neighbours are loaded from a valid address (the centre) and
// replaced by zero with a select, so every load is unconditional and the
// whole next plane is in flight while the current one is computed.
//
// Tiles (gx k-tiles x gy j-tiles x gz i-chunks, k fastest) are dealt to the 8
// XCDs in contiguous runs: the hardware hands blocks to the XCDs round robin,
// so block b takes tile (b % 8) * (G / 8) + b / 8 of a 1-D grid of G (a
// multiple of 8) blocks.  The tiles an XCD runs at the same time are then
// neighbours in j and k and their halo rows are L2 hits (with the natural
// order, j-neighbouring tiles sit on different XCDs and every halo row is
// fetched from HBM again).  A reduction's partials keep the natural tile
// order, so sums are the same as with a 3-D grid.
template <int TK, bool REDUCE, bool MARCH, class Load, class Epi>
__device__ inline void stencil_body(GridPos g, int n, int ichunk, int gx, int gy, int gz,
                                    const Load& ldr, const Epi& epi, const Reduce& red) {
    constexpr int TJ = 256 / TK;
    const unsigned G8 = g.nb >> 3;
    const unsigned lin = (g.bid & 7u) * G8 + (g.bid >> 3);
    if (lin >= (unsigned)(gx * gy * gz)) return;  // padding block, uniform
    const int bx = (int)(lin % (unsigned)gx), by = (int)((lin / (unsigned)gx) % (unsigned)gy);
    const int bz = (int)(lin / (unsigned)(gx * gy));
    const Load L = ldr.prep();
    const Epi E = epi.prep();
    using P = decltype(L.make(L.load(int64_t(0))));
    using X = decltype(E.preload(int64_t(0)));
    const int tk = threadIdx.x % TK, tj = threadIdx.x / TK;
    const int k = bx * TK + tk, j = by * TJ + tj;
    const int i0 = bz * ichunk;
    const int i1 = min(n, i0 + ichunk);
    double acc = 0.0;
    if (k < n && j < n && i0 < n) {
        const int64_t N1 = n, N2 = N1 * N1;
        int64_t id = ((int64_t)i0 * N1 + j) * N1 + k;
        const bool hj0 = j > 0, hj1 = j + 1 < n, hk0 = k > 0, hk1 = k + 1 < n;
        auto fetch = [&](int64_t idx, int64_t safe, bool ok) -> P {
            const P v = L.make(L.load(ok ? idx : safe));
            return ok ? v : P::zero();
        };
        P pm = fetch(id - N2, id, i0 > 0);
        P pc = L.make(L.load(id));
        P pp = fetch(id + N2, id, i0 + 1 < n);
        P jm = fetch(id - N1, id, hj0);
        P jp = fetch(id + N1, id, hj1);
        P km = fetch(id - 1, id, hk0);
        P kp = fetch(id + 1, id, hk1);
        X ex = E.preload(id);
        if constexpr (!MARCH) {
            acc += E(id, ex, pc, pm, pp, jm, jp, km, kp);
        } else {
            for (int i = i0; i < i1; ++i, id += N2) {
                // prefetch plane i+1 (its centre pp is already resident)
                const bool more = i + 1 < i1;
                const int64_t nx = more ? id + N2 : id;
                const P qp = fetch(nx + N2, nx, more && i + 2 < n);
                const P qjm = fetch(nx - N1, nx, hj0);
                const P qjp = fetch(nx + N1, nx, hj1);
                const P qkm = fetch(nx - 1, nx, hk0);
                const P qkp = fetch(nx + 1, nx, hk1);
                const X qex = E.preload(nx);
                acc += E(id, ex, pc, pm, pp, jm, jp, km, kp);
                pm = pc;
                pc = pp;
                pp = qp;
                jm = qjm;
                jp = qjp;
                km = qkm;
                kp = qkp;
                ex = qex;
            }
        }
    }
    if constexpr (REDUCE) block_partial<256>(acc, red, lin);
}
template <int TK, bool REDUCE, bool MARCH, class Load, class Epi>
__global__ __launch_bounds__(256) void k_stencil(int n, int ichunk, int gx, int gy, int gz,
                                                  Load ldr, Epi epi, Reduce red) {
    stencil_body<TK, REDUCE, MARCH, Load, Epi>(hw_grid(), n, ichunk, gx, gy, gz, ldr, epi, red);
}

// LDS-tiled form for large grids.  A block of 256 threads owns a 64 (k) x
// 4*RW (j) tile (rows tj, tj + 4, ... of each thread) and a chunk of planes,
// and marches along i loading every node once: the own nodes' loads run three
// planes ahead of the plane being computed and the halo ring's two (the ring
// of 2 rows of 64 and 2 columns of 4*RW nodes is loaded by the first threads,
// one node each), so two planes of HBM traffic are in flight per block.  The
// plane's values go to a double-buffered LDS plane with the ring, where the
// j / k neighbours are read; the i neighbours stay in registers.  One barrier
// per plane.  Same Loader / Epilogue interface, arithmetic and outputs as
// k_stencil.  A reduction writes one partial per block, or with PLANES one
// partial per plane and group of 4 rows: the partials of the one-node-per-
// thread k_stencil (64 x 4 tiles, one plane each), in its tile order and with
// its summation order (the same 64-lane wave sums, added over the 4 waves in
// order), so the reduced total is bit for bit k_stencil's.
template <int RW, bool REDUCE, bool PLANES, class Load, class Epi>
__device__ inline void stencil_lds_body(GridPos g, int n, int ichunk, int gx, int gy, int gz,
                                        const Load& ldr, const Epi& epi, const Reduce& red) {
    constexpr int TK = 64, TJ = 4 * RW;
    const unsigned G8 = g.nb >> 3;
    const unsigned lin = (g.bid & 7u) * G8 + (g.bid >> 3);
    if (lin >= (unsigned)(gx * gy * gz)) return;  // padding block, uniform
    const int bx = (int)(lin % (unsigned)gx), by = (int)((lin / (unsigned)gx) % (unsigned)gy);
    const int bz = (int)(lin / (unsigned)(gx * gy));
    const Load L = ldr.prep();
    const Epi E = epi.prep();
    using R = decltype(L.load(int64_t(0)));
    using P = decltype(L.make(L.load(int64_t(0))));
    using X = decltype(E.preload(int64_t(0)));
    __shared__ P buf[2][TJ + 2][TK + 2];

    const int tk = threadIdx.x & 63, tj = threadIdx.x >> 6;
    const int k0 = bx * TK, j0 = by * TJ, k = k0 + tk;
    const int i0 = bz * ichunk, i1 = min(n, i0 + ichunk);
    const int64_t N1 = n, N2 = N1 * N1;
    // node (q, j, k) is id = q N2 + col with col = j n + k < 2^22 (n <= 2048):
    // the plane part is uniform, col a 32-bit per-thread offset
    bool own[RW];
    uint32_t col[RW];  // 0 for columns outside the domain (a valid address)
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const int j = j0 + tj + 4 * r;
        own[r] = j < n && k < n;
        col[r] = own[r] ? (uint32_t)(j * n + k) : 0u;
    }
    // halo node of this thread: threads 0..127 the rows j0-1 and j0+TJ,
    // threads 128..128+2TJ-1 the columns k0-1 and k0+TK
    int hr = -1, hc = 0, hj = -1, hk = -1;
    if (threadIdx.x < 128) {
        hr = threadIdx.x < 64 ? 0 : TJ + 1;
        hc = tk + 1;
        hj = threadIdx.x < 64 ? j0 - 1 : j0 + TJ;
        hk = k;
    } else if (threadIdx.x < 128 + 2 * TJ) {
        const int t = threadIdx.x - 128, side = t / TJ;
        hr = t % TJ + 1;
        hc = side ? TK + 1 : 0;
        hj = j0 + t % TJ;
        hk = side ? k0 + TK : k0 - 1;
    }
    const bool hown = hr >= 0 && hj >= 0 && hj < n && hk >= 0 && hk < n;
    const uint32_t hcol = hown ? (uint32_t)(hj * n + hk) : 0u;
    // loads of planes outside the domain go to plane 0, of columns outside it
    // to offset 0; their values are replaced by zero
    auto plane_of = [&](int q) -> int64_t { return (unsigned)q < (unsigned)n ? q * N2 : 0; };
    auto own_ok = [&](int r, int q) { return own[r] && (unsigned)q < (unsigned)n; };

    // Register queues indexed by a plane's position t = p - i0 in the chunk
    // modulo 3 (P: planes i-1, i, i+1; Q: raw own values of planes i+1..i+3;
    // E: epilogue loads of planes i..i+2; H: halo raw values of planes
    // i..i+2), the LDS plane by t mod 2.  The march is unrolled by 6 planes
    // so that every slot is a compile-time register (no values move).
    P Pv[3][RW];
    R Q[3][RW];
    X Ev[3][RW];
    R H[3];
    H[0] = L.load(plane_of(i0) + hcol);
    H[1] = L.load(plane_of(i0 + 1) + hcol);
    H[2] = H[1];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
        const R a = L.load(plane_of(i0 - 1) + col[r]), b = L.load(plane_of(i0) + col[r]);
        Q[1][r] = L.load(plane_of(i0 + 1) + col[r]);
        Q[2][r] = L.load(plane_of(i0 + 2) + col[r]);
        Q[0][r] = Q[2][r];
        Ev[0][r] = E.preload(plane_of(i0) + col[r]);
        Ev[1][r] = E.preload(plane_of(i0 + 1) + col[r]);
        Ev[2][r] = Ev[1][r];
        Pv[2][r] = own_ok(r, i0 - 1) ? L.make(a) : P::zero();
        Pv[0][r] = own_ok(r, i0) ? L.make(b) : P::zero();
        Pv[1][r] = Pv[0][r];
    }
    double acc = 0.0;
    // PLANES: the wave sums of plane i, row group r, in slot i mod 2 (written
    // after the plane's barrier, added up by thread 0 after the next one)
    __shared__ double wsum[PLANES && REDUCE ? 2 : 1][PLANES && REDUCE ? RW : 1][4];
    const int gy4 = (n + 3) / 4;  // row groups of k_stencil
    auto flush = [&](int slot, int i) {
        if constexpr (PLANES && REDUCE) {
            if (threadIdx.x == 0) {
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const int g4 = by * RW + r;
                    if (g4 >= gy4) continue;  // rows past n: no such k_stencil tile
                    double t = 0.0;
                    for (int w = 0; w < 4; ++w) t += wsum[slot][r][w];
                    red.partials[(unsigned)bx + (unsigned)gx * ((unsigned)g4 + (unsigned)gy4 * (unsigned)i)] = t;
                }
            }
        }
    };
    // plane i = i0 + t, PH = t mod 6
    auto plane = [&](auto ph, int i) {
        constexpr int PH = decltype(ph)::value;
        constexpr int S0 = PH % 3, S1 = (PH + 1) % 3, S2 = (PH + 2) % 3;  // t, t+1, t+2 (= t-1)
        P(&B)[TJ + 2][TK + 2] = buf[PH & 1];
#pragma unroll
        for (int r = 0; r < RW; ++r) B[tj + 4 * r + 1][tk + 1] = Pv[S0][r];
        if (hr >= 0) B[hr][hc] = hown ? L.make(H[S0]) : P::zero();
        if (i + 3 <= i1) {  // plane i+3 into the slot of plane i (consumed)
#pragma unroll
            for (int r = 0; r < RW; ++r) Q[S0][r] = L.load(plane_of(i + 3) + col[r]);
        }
        if (i + 2 < i1) {  // plane i+2 into the slots of plane i-1
            const int64_t pb = plane_of(i + 2);
#pragma unroll
            for (int r = 0; r < RW; ++r) Ev[S2][r] = E.preload(pb + col[r]);
            H[S2] = L.load(pb + hcol);
        }
        __syncthreads();
        if constexpr (PLANES && REDUCE) {
            if (i > i0) flush((PH + 1) & 1, i - 1);
        }
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const P pp = own_ok(r, i + 1) ? L.make(Q[S1][r]) : P::zero();
            const int row = tj + 4 * r + 1;
            if constexpr (PLANES && REDUCE) {
                double v = 0.0;  // as k_stencil's per-thread sum
                if (own[r])
                    v += E((int64_t)i * N2 + col[r], Ev[S0][r], Pv[S0][r], Pv[S2][r], pp,
                           B[row - 1][tk + 1], B[row + 1][tk + 1], B[row][tk], B[row][tk + 2]);
                const double ws = wave_sum(v);
                if (tk == 0) wsum[PH & 1][r][tj] = ws;
            } else if (own[r]) {
                acc += E((int64_t)i * N2 + col[r], Ev[S0][r], Pv[S0][r], Pv[S2][r], pp,
                         B[row - 1][tk + 1], B[row + 1][tk + 1], B[row][tk], B[row][tk + 2]);
            }
            Pv[S1][r] = pp;
        }
    };
    int t0 = 0;
    const int count = i1 - i0;
    for (; t0 + 6 <= count; t0 += 6)
        static_for<6>([&](auto ph) { plane(ph, i0 + t0 + decltype(ph)::value); });
    static_for<6>([&](auto ph) {
        constexpr int PH = decltype(ph)::value;
        if (t0 + PH < count) plane(ph, i0 + t0 + PH);
    });
    if constexpr (REDUCE && PLANES) {
        __syncthreads();
        if (count > 0) flush((count - 1) & 1, i1 - 1);
    } else if constexpr (REDUCE) {
        block_partial<256>(acc, red, lin);
    }
}
template <int RW, bool REDUCE, bool PLANES, class Load, class Epi>
__global__ __launch_bounds__(256) void k_stencil_lds(int n, int ichunk, int gx, int gy, int gz,
                                                      Load ldr, Epi epi, Reduce red) {
    stencil_lds_body<RW, REDUCE, PLANES, Load, Epi>(hw_grid(), n, ichunk, gx, gy, gz, ldr, epi,
                                                    red);
}

// Launch shape: 64x4 (k x j) tiles for n > 32, 32x8 below.  Epilogues with an
// FP64 exp() per node (Epi::kMarch) march along i in chunks of 8 planes on
// large grids (register queue, next plane prefetched); the others use one node
// per thread, which streams best for cheap epilogues.  At most kMaxPartials
// blocks.
struct StencilShape {
    int tk, ichunk;
    bool march;
    int gx, gy, gz;   // tiles
    unsigned blocks;  // launched blocks: tiles rounded up to a multiple of 8
};

inline StencilShape stencil_shape(int64_t n, bool heavy) {
    StencilShape s;
    s.tk = n > 32 ? 64 : 32;
    const int tj = 256 / s.tk;
    const int64_t gx = (n + s.tk - 1) / s.tk, gy = (n + tj - 1) / tj;
    int ich = (heavy && n >= 128) ? 8 : 1;
    while (gx * gy * ((n + ich - 1) / ich) > kMaxPartials) ich *= 2;
    s.ichunk = ich;
    s.march = ich > 1;
    s.gx = (int)gx;
    s.gy = (int)gy;
    s.gz = (int)((n + ich - 1) / ich);
    const int64_t tiles = gx * gy * (int64_t)s.gz;
    s.blocks = (unsigned)((tiles + 7) / 8 * 8);
    return s;
}

// Grids from kLdsMinN up take the LDS-tiled form: 64 x 8 tiles, and the plane
// chunk that fills the device in whole waves of resident blocks (fill_chunk,
// with the kernel's occupancy).
constexpr int64_t kLdsMinN = 128;
constexpr int kLdsRows = 2;

inline StencilShape stencil_shape_lds(int64_t n, int64_t slots) {
    StencilShape s;
    s.tk = 64;
    const int64_t gx = (n + 63) / 64, gy = (n + 4 * kLdsRows - 1) / (4 * kLdsRows);
    int ich = fill_chunk(n, gx * gy, slots, 2, 4, 64);
    while (gx * gy * ((n + ich - 1) / ich) > kMaxPartials) ich *= 2;
    s.ichunk = ich;
    s.march = true;
    s.gx = (int)gx;
    s.gy = (int)gy;
    s.gz = (int)((n + ich - 1) / ich);
    const int64_t tiles = gx * gy * (int64_t)s.gz;
    s.blocks = (unsigned)((tiles + 7) / 8 * 8);
    return s;
}

// Grids from kPlanesMinN to kLdsMinN take the LDS-tiled form with the
// one-node-per-thread form's reduction partials (the same values, sums and
// totals as before, bit for bit).
constexpr int64_t kPlanesMinN = 33;

// Reduction partials of a launch over an n^3 grid: k_stencil's tiles (64 x 4
// x 1 for n > 32), which the LDS-tiled form also writes below kLdsMinN.
inline unsigned stencil_partials(int64_t n, bool heavy) {
    if (n >= kLdsMinN) return 0;  // the LDS-tiled form's own count (its launch shape)
    const StencilShape s = stencil_shape(n, heavy);
    return (unsigned)(s.gx * s.gy * s.gz);
}

template <bool REDUCE, class Load, class Epi>
void launch_stencil(Ctx& c, int64_t n, const Load& ldr, const Epi& epi, const Epilogue& ep) {
    if (n >= kLdsMinN) {
        static const int per_cu =
            blocks_per_cu(k_stencil_lds<kLdsRows, REDUCE, false, Load, Epi>, 256);
        const StencilShape s = stencil_shape_lds(n, (int64_t)per_cu * c.cu_count);
        hipLaunchKernelGGL((k_stencil_lds<kLdsRows, REDUCE, false, Load, Epi>), dim3(s.blocks),
                           dim3(256), 0, c.stream, (int)n, s.ichunk, s.gx, s.gy, s.gz, ldr, epi,
                           c.red);
        NUMA_HIP_LAUNCH_CHECK();
        if constexpr (REDUCE) c.finish((unsigned)(s.gx * s.gy * s.gz), ep);
        return;
    }
    if (n >= kPlanesMinN) {
        static const int per_cu =
            blocks_per_cu(k_stencil_lds<kLdsRows, REDUCE, true, Load, Epi>, 256);
        const StencilShape s = stencil_shape_lds(n, (int64_t)per_cu * c.cu_count);
        hipLaunchKernelGGL((k_stencil_lds<kLdsRows, REDUCE, true, Load, Epi>), dim3(s.blocks),
                           dim3(256), 0, c.stream, (int)n, s.ichunk, s.gx, s.gy, s.gz, ldr, epi,
                           c.red);
        NUMA_HIP_LAUNCH_CHECK();
        if constexpr (REDUCE) c.finish(stencil_partials(n, Epi::kMarch), ep);
        return;
    }
    const StencilShape s = stencil_shape(n, Epi::kMarch);
#define NUMA_STENCIL_LAUNCH(TK_, MARCH_)                                                     \
    hipLaunchKernelGGL((k_stencil<TK_, REDUCE, MARCH_, Load, Epi>), dim3(s.blocks), dim3(256), \
                       0, c.stream, (int)n, s.ichunk, s.gx, s.gy, s.gz, ldr, epi, c.red)
    if (s.tk == 64) {
        if (s.march)
            NUMA_STENCIL_LAUNCH(64, true);
        else
            NUMA_STENCIL_LAUNCH(64, false);
    } else {
        if (s.march)
            NUMA_STENCIL_LAUNCH(32, true);
        else
            NUMA_STENCIL_LAUNCH(32, false);
    }
#undef NUMA_STENCIL_LAUNCH
    NUMA_HIP_LAUNCH_CHECK();
    if constexpr (REDUCE) c.finish((unsigned)(s.gx * s.gy * s.gz), ep);
}

// ---------------------------------------------------------------------------
// Finite-difference step h = root * (1 + ||x||) / ||v|| with ||x|| and ||v||
// taken from device scalar slots, so a Krylov step never waits for the host
// (operator.hpp:107-108 with root = sqrt(eps); :124-125 with root = eps^(1/4)).
// ---------------------------------------------------------------------------
struct FDStep {
    double root;
    const double* xnorm;  // ||x||
    const double* vnorm;  // ||v||
    __device__ double h() const {
#pragma clang fp contract(off)
        const double nv = *vnorm;
        return nv == 0.0 ? 0.0 : root * (1.0 + *xnorm) / nv;
    }
};

// ---------------------------------------------------------------------------
// Loaders
// ---------------------------------------------------------------------------
struct NoExtra {};

template <class T>
struct LdVec {  // w = v
    const T* v;
    __device__ LdVec prep() const { return *this; }
    __device__ T load(int64_t id) const { return v[id]; }
    __device__ P1 make(T r) const { return {static_cast<double>(r)}; }
};

// w = (x - a z) - b u   (Newton trial x - lambda z, krylov_newton.hpp:99;
// Chebyshev step x - z - u/2, :166); HZ / HU select the terms at compile time.
template <class TZ, bool HZ, bool HU>
struct LdStep {
    const double* x;
    const TZ* z;
    double a;
    const double* u;
    double b;
    struct Raw {
        double x;
        TZ z;
        double u;
    };
    __device__ LdStep prep() const { return *this; }
    __device__ Raw load(int64_t id) const {
        Raw r{x[id], TZ(0), 0.0};
        if constexpr (HZ) r.z = z[id];
        if constexpr (HU) r.u = u[id];
        return r;
    }
    __device__ P1 make(const Raw& r) const {
#pragma clang fp contract(off)
        double w = r.x;
        if constexpr (HZ) w = w - a * static_cast<double>(r.z);
        if constexpr (HU) w = w - b * r.u;
        return {w};
    }
};

// w = x + h v (operator.hpp:109-110).
template <class TV>
struct LdXpHV {
    const double* x;
    const TV* v;
    FDStep step;
    double h;
    struct Raw {
        double x;
        TV v;
    };
    __device__ LdXpHV prep() const {
        LdXpHV r = *this;
        r.h = step.h();
        return r;
    }
    __device__ Raw load(int64_t id) const { return {x[id], v[id]}; }
    __device__ P1 make(const Raw& r) const {
#pragma clang fp contract(off)
        return {r.x + h * static_cast<double>(r.v)};
    }
};

// (x + h v, x + (-h) v) for the central difference (operator.hpp:126-128).
template <class TV>
struct LdPM {
    const double* x;
    const TV* v;
    FDStep step;
    double h;
    struct Raw {
        double x;
        TV v;
    };
    __device__ LdPM prep() const {
        LdPM r = *this;
        r.h = step.h();
        return r;
    }
    __device__ Raw load(int64_t id) const { return {x[id], v[id]}; }
    __device__ P2 make(const Raw& r) const {
#pragma clang fp contract(off)
        const double vv = static_cast<double>(r.v);
        return {r.x + h * vv, r.x + (-h) * vv};
    }
};

// ---------------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------------
// Every Bratu epilogue takes lambda by value or, when lam_ptr is set, from a
// device scalar read once per thread in prep() (a solver's lambda, so that a
// captured graph stays valid when lambda changes).
//
// F = h2inv*(6w - sum nbrs) - lambda e^w (bratu3d.hpp:21-31); with WOUT also
// writes the trial point w itself; returns F^2.
template <bool WOUT>
struct EpF {
    static constexpr bool kMarch = true;
    double h2inv, lambda;
    double* F;
    double* wout;
    const double* lam_ptr = nullptr;
    __device__ EpF prep() const {
        EpF r = *this;
        if (lam_ptr) r.lambda = *lam_ptr;
        return r;
    }
    __device__ NoExtra preload(int64_t) const { return {}; }
    __device__ double operator()(int64_t id, NoExtra, P1 c, P1 im, P1 ip, P1 jm, P1 jp, P1 km,
                                 P1 kp) const {
#pragma clang fp contract(off)
        double f = h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a);
        f -= lambda * exp(c.a);
        F[id] = f;
        if constexpr (WOUT) wout[id] = c.a;
        return f * f;
    }
};

// J v epilogue options: DOT reduces out*dotv[id] (first MGS dot), RESID writes
// b - J v and reduces its square.
template <class TO>
struct JvExtra {
    double f0;   // F0 (FD), x (analytic) or lambda e^x (rearranged FD)
    TO dv;
    double b;
};

// Forward-difference J v = (F(x+hv) - F0)/h (operator.hpp:113), multiplied by
// the Arnoldi scale *scale_ptr (1 if null).
template <class TO, bool DOT, bool RESID>
struct EpJvFD {
    static constexpr bool kMarch = true;
    double h2inv, lambda;
    FDStep step;
    const double* scale_ptr;
    const double* F0;
    TO* out;
    const TO* dotv;
    const double* b;
    double h, scale;
    const double* lam_ptr = nullptr;
    __device__ EpJvFD prep() const {
        EpJvFD r = *this;
        r.h = step.h();
        r.scale = scale_ptr ? *scale_ptr : 1.0;
        if (lam_ptr) r.lambda = *lam_ptr;
        return r;
    }
    __device__ JvExtra<TO> preload(int64_t id) const {
        JvExtra<TO> e{F0[id], TO(0), 0.0};
        if constexpr (DOT) e.dv = dotv[id];
        if constexpr (RESID) e.b = b[id];
        return e;
    }
    __device__ double operator()(int64_t id, const JvExtra<TO>& e, P1 c, P1 im, P1 ip, P1 jm,
                                 P1 jp, P1 km, P1 kp) const {
#pragma clang fp contract(off)
        double val = 0.0;
        if (h != 0.0) {  // ||v|| == 0 gives J v = 0 (operator.hpp:105-106)
            double f1 = h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a);
            f1 -= lambda * exp(c.a);
            val = (f1 - e.f0) / h;
            val = val * scale;
        }
        if constexpr (RESID) val = e.b - val;
        const TO o = static_cast<TO>(val);
        out[id] = o;
        if constexpr (DOT) return static_cast<double>(o) * static_cast<double>(e.dv);
        if constexpr (RESID) return static_cast<double>(o) * static_cast<double>(o);
        return 0.0;
    }
};

// (e^{h v} - 1) / h in FP64 as v g(y), y = h v, g(y) = (e^y - 1) / y.  For
// |y| <= 2^-10, g is its Taylor polynomial to y^4 / 120 (the first omitted
// term, y^5 / 720, is below 2^-59 relative): 4 FMAs instead of an FP64 expm1
// and an FP64 division.  The forward-difference step h = sqrt(eps) (1 + ||x||)
// / ||v|| keeps |y| <= sqrt(eps) (1 + ||x||) there for any ||x|| below 6e4;
// beyond, expm1(y) / y.
__device__ __forceinline__ double expm1_quotient(double h, double v) {
    const double y = h * v;
    double g;
    if (fabs(y) <= 0x1p-10) {
        g = __builtin_fma(y, 1.0 / 120.0, 1.0 / 24.0);
        g = __builtin_fma(y, g, 1.0 / 6.0);
        g = __builtin_fma(y, g, 0.5);
        g = __builtin_fma(y, g, 1.0);
    } else {
        g = expm1(y) / y;
    }
    return v * g;
}

// The same forward-difference quotient for the mixed-precision solver, in
// the rearranged form
//   (F(x + h v) - F(x)) / h = h2inv (6v - sum nbrs) - lambda e^x (e^{h v} - 1) / h
// (the Laplacian is linear, so only the exponential is differenced, by
// expm1_quotient in FP64).  It is the quotient of EpJvFD without the rounding
// of x + h v and of the subtraction F(x + h v) - F(x), and it reads v and
// E = lambda e^x (FP32, written once per Newton point) instead of x and F(x)
// in FP64: 16 instead of 28 bytes per node with the fused first MGS dot.
// The loader is LdVec over v.
template <class TO, bool DOT, bool RESID>
struct EpJvFDe {
    static constexpr bool kMarch = true;
    double h2inv;
    FDStep step;
    const double* scale_ptr;
    const float* E;
    TO* out;
    const TO* dotv;
    const double* b;
    double h, scale;
    __device__ EpJvFDe prep() const {
        EpJvFDe r = *this;
        r.h = step.h();
        r.scale = scale_ptr ? *scale_ptr : 1.0;
        return r;
    }
    __device__ JvExtra<TO> preload(int64_t id) const {
        JvExtra<TO> e{static_cast<double>(E[id]), TO(0), 0.0};
        if constexpr (DOT) e.dv = dotv[id];
        if constexpr (RESID) e.b = b[id];
        return e;
    }
    __device__ double operator()(int64_t id, const JvExtra<TO>& e, P1 c, P1 im, P1 ip, P1 jm,
                                 P1 jp, P1 km, P1 kp) const {
#pragma clang fp contract(off)
        double val = 0.0;
        if (h != 0.0) {  // ||v|| == 0 gives J v = 0 (operator.hpp:105-106)
            val = h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a);
            val -= e.f0 * expm1_quotient(h, c.a);
            val = val * scale;
        }
        if constexpr (RESID) val = e.b - val;
        const TO o = static_cast<TO>(val);
        out[id] = o;
        if constexpr (DOT) return static_cast<double>(o) * static_cast<double>(e.dv);
        if constexpr (RESID) return static_cast<double>(o) * static_cast<double>(o);
        return 0.0;
    }
};

// Analytic J v = h2inv*(6v - sum nbrs) - lambda e^x v (bratu3d.hpp:33-43).
template <class TO, bool DOT, bool RESID>
struct EpJvAn {
    static constexpr bool kMarch = true;
    double h2inv, lambda;
    const double* scale_ptr;
    const double* x;
    TO* out;
    const TO* dotv;
    const double* b;
    double scale;
    const double* lam_ptr = nullptr;
    __device__ EpJvAn prep() const {
        EpJvAn r = *this;
        r.scale = scale_ptr ? *scale_ptr : 1.0;
        if (lam_ptr) r.lambda = *lam_ptr;
        return r;
    }
    __device__ JvExtra<TO> preload(int64_t id) const {
        JvExtra<TO> e{x[id], TO(0), 0.0};
        if constexpr (DOT) e.dv = dotv[id];
        if constexpr (RESID) e.b = b[id];
        return e;
    }
    __device__ double operator()(int64_t id, const JvExtra<TO>& e, P1 c, P1 im, P1 ip, P1 jm,
                                 P1 jp, P1 km, P1 kp) const {
#pragma clang fp contract(off)
        double val = h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a);
        val -= lambda * exp(e.f0) * c.a;
        val = val * scale;
        if constexpr (RESID) val = e.b - val;
        const TO o = static_cast<TO>(val);
        out[id] = o;
        if constexpr (DOT) return static_cast<double>(o) * static_cast<double>(e.dv);
        if constexpr (RESID) return static_cast<double>(o) * static_cast<double>(o);
        return 0.0;
    }
};

// Central-difference F''[v,v] = (F(x+hv) - 2F0 + F(x-hv))/h^2 (operator.hpp:132-133).
struct EpD2FD {
    static constexpr bool kMarch = true;
    double h2inv, lambda;
    FDStep step;
    const double* F0;
    double* out;
    double h;
    const double* lam_ptr = nullptr;
    __device__ EpD2FD prep() const {
        EpD2FD r = *this;
        r.h = step.h();
        if (lam_ptr) r.lambda = *lam_ptr;
        return r;
    }
    __device__ double preload(int64_t id) const { return F0[id]; }
    __device__ double operator()(int64_t id, double f0, P2 c, P2 im, P2 ip, P2 jm, P2 jp, P2 km,
                                 P2 kp) const {
#pragma clang fp contract(off)
        if (h == 0.0) {
            out[id] = 0.0;
            return 0.0;
        }
        double fp = h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a);
        fp -= lambda * exp(c.a);
        double fm = h2inv * lap_num(c.b, im.b, ip.b, jm.b, jp.b, km.b, kp.b);
        fm -= lambda * exp(c.b);
        out[id] = (fp - 2.0 * f0 + fm) / (h * h);
        return 0.0;
    }
};

// Plain -Lap_h v.
template <class TO>
struct EpLap {
    static constexpr bool kMarch = false;
    double h2inv;
    TO* out;
    __device__ EpLap prep() const { return *this; }
    __device__ NoExtra preload(int64_t) const { return {}; }
    __device__ double operator()(int64_t id, NoExtra, P1 c, P1 im, P1 ip, P1 jm, P1 jp, P1 km,
                                 P1 kp) const {
#pragma clang fp contract(off)
        out[id] = static_cast<TO>(h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a));
        return 0.0;
    }
};

}  // namespace numa_hip
