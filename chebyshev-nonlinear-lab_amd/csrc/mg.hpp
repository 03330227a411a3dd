// mg.hpp - geometric multigrid V-cycle for -Lap_h (7-point, zero Dirichlet)
// on gfx950.  Restates linalg/poisson_mg.hpp:24-320 (RB-GS smoother, 27-point
// full-weighting restriction with coarse node I <-> fine node 2I+1, trilinear
// prolongation, coarse h^-2 /4 per level) and extends it with:
//   * even-n coarsening n -> n/2 with the same transfer operators (the
//     reference stops at the first even n, poisson_mg.hpp:60-62);
//   * weighted-Jacobi and Chebyshev (degree d, interval [lmax/ratio, lmax] of
//     D^-1 A with the exact lmax = 1 + cos(pi/(n+1))) smoothers;
//   * a Chebyshev semi-iteration over the exact spectrum
//     [1 - cos(pi/(n+1)), 1 + cos(pi/(n+1))] as coarse solve, run by one
//     workgroup with the level resident in LDS.
// Level storage is FP64 or FP32 (T).  The polynomial smoothing steps compute in
// the storage precision (SmoothArith<T>); the transfers and the coarse solve
// compute in FP64 per point.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <utility>

#include <hip/hip_ext.h>

#include "blas.hpp"
#include "smooth_fused.hpp"
#include "stencil.hpp"
#include "tune.hpp"

namespace numa_hip {

// One Jacobi-type / Chebyshev step:  d <- c1 d + c2 (b - A x);  xn = x + d,
// in the arithmetic SmoothArith<T> (smooth_fused.hpp).
// USE_D reads the previous direction, WRITE_D stores the new one.
template <class T, bool USE_D, bool WRITE_D>
struct EpPoly {
    using A = SmoothArith<T>;
    static constexpr bool kMarch = false;
    double h2inv, c1, c2;
    const T* b;
    T* d;
    T* xn;
    struct Extra {
        T b, d;
    };
    __device__ EpPoly prep() const { return *this; }
    __device__ Extra preload(int64_t id) const {
        Extra e{b[id], T(0)};
        if constexpr (USE_D) e.d = d[id];
        return e;
    }
    __device__ double operator()(int64_t id, const Extra& e, P1 c, P1 im, P1 ip, P1 jm, P1 jp,
                                 P1 km, P1 kp) const {
#pragma clang fp contract(off)
        const A cc = static_cast<A>(c.a);
        const A res = poly_res<A>(static_cast<A>(e.b), static_cast<A>(h2inv), cc,
                                  static_cast<A>(im.a), static_cast<A>(ip.a),
                                  static_cast<A>(jm.a), static_cast<A>(jp.a),
                                  static_cast<A>(km.a), static_cast<A>(kp.a));
        const A dn = USE_D ? poly_dir<A>(static_cast<A>(c1), static_cast<A>(e.d),
                                         static_cast<A>(c2), res)
                           : static_cast<A>(c2) * res;
        if constexpr (WRITE_D) d[id] = static_cast<T>(dn);
        xn[id] = static_cast<T>(cc + dn);
        return 0.0;
    }
};

// r = b - A x   (poisson_mg.hpp:128-149)
template <class T>
struct EpResid {
    static constexpr bool kMarch = false;
    double h2inv;
    const T* b;
    T* r;
    __device__ EpResid prep() const { return *this; }
    __device__ T preload(int64_t id) const { return b[id]; }
    __device__ double operator()(int64_t id, T bv, P1 c, P1 im, P1 ip, P1 jm, P1 jp, P1 km,
                                 P1 kp) const {
#pragma clang fp contract(off)
        r[id] = static_cast<T>(static_cast<double>(bv) -
                               h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a));
        return 0.0;
    }
};

// First step from x = 0:  d = c2 b;  xn = d  (arithmetic SmoothArith<T>).
template <bool VEC, class T>
__device__ inline void poly_first_body(GridPos g, int64_t N, double c2d, const T* b, T* d, T* xn) {
#pragma clang fp contract(off)
    using A = SmoothArith<T>;
    const A c2 = static_cast<A>(c2d);
    auto f = [c2](Pair<T> v) {
        return Pair<T>{static_cast<T>(c2 * (A)v.a), static_cast<T>(c2 * (A)v.b)};
    };
    for_each_pair(
        g, N,
        [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
            Pair<T> b0{}, b1{};
            if (ok0) b0 = ldp<VEC>(b, q0);
            if (ok1) b1 = ldp<VEC>(b, q1);
            const Pair<T> o0 = f(b0), o1 = f(b1);
            if (ok0) {
                if (d) stp<VEC>(d, q0, o0);
                stp<VEC>(xn, q0, o0);
            }
            if (ok1) {
                if (d) stp<VEC>(d, q1, o1);
                stp<VEC>(xn, q1, o1);
            }
        },
        [&](int64_t i) {
            const A dn = c2 * static_cast<A>(b[i]);
            if (d) d[i] = static_cast<T>(dn);
            xn[i] = static_cast<T>(dn);
        });
}
template <bool VEC, class T>
__global__ __launch_bounds__(kBlock) void k_poly_first(int64_t N, double c2d, const T* b, T* d,
                                                        T* xn) {
    poly_first_body<VEC, T>(hw_grid(), N, c2d, b, d, xn);
}

// Red-black Gauss-Seidel half sweep on one colour (poisson_mg.hpp:107-124):
// x_id = (b_id/h2inv + sum nbrs)/6 for (i+j+k) % 2 == color, in place.
template <class T>
__global__ __launch_bounds__(kBlock) void k_rbgs(int n, double inv, int color, const T* b, T* x) {
#pragma clang fp contract(off)
    const int64_t N1 = n, N2 = N1 * N1, N = N2 * N1;
    for (int64_t id = (int64_t)blockIdx.x * kBlock + threadIdx.x; id < N;
         id += (int64_t)gridDim.x * kBlock) {
        const int k = (int)(id % N1);
        const int j = (int)((id / N1) % N1);
        const int i = (int)(id / N2);
        if (((i + j + k) & 1) != color) continue;
        // the seven loads are issued before the first use (a neighbour outside
        // the domain loads the node itself and is not added)
        const bool im = i > 0, ip = i + 1 < n, jm = j > 0, jp = j + 1 < n, km = k > 0,
                   kp = k + 1 < n;
        const T vb = b[id];
        const T vim = x[im ? id - N2 : id], vip = x[ip ? id + N2 : id];
        const T vjm = x[jm ? id - N1 : id], vjp = x[jp ? id + N1 : id];
        const T vkm = x[km ? id - 1 : id], vkp = x[kp ? id + 1 : id];
        double s = static_cast<double>(vb) * inv;
        if (im) s += static_cast<double>(vim);
        if (ip) s += static_cast<double>(vip);
        if (jm) s += static_cast<double>(vjm);
        if (jp) s += static_cast<double>(vjp);
        if (km) s += static_cast<double>(vkm);
        if (kp) s += static_cast<double>(vkp);
        x[id] = static_cast<T>(s / 6.0);
    }
}

// 27-point full weighting, coarse I <-> fine 2I+1 (poisson_mg.hpp:152-188).
// One thread per coarse node, 64 (K) x 4 (J) threads per block, grid z = I.
template <class T>
__global__ __launch_bounds__(256) void k_restrict(int nf, int nc, const T* r, T* bc) {
#pragma clang fp contract(off)
    const double w[3] = {0.25, 0.5, 0.25};
    const int K = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int I = blockIdx.z;
    if (K >= nc || J >= nc) return;
    const int64_t F1 = nf, F2 = F1 * F1;
    double sum = 0.0;
    for (int a = 0; a < 3; ++a) {
        const int fi = 2 * I + a;
        if (fi >= nf) continue;
        for (int bb = 0; bb < 3; ++bb) {
            const int fj = 2 * J + bb;
            if (fj >= nf) continue;
            const int64_t row = (int64_t)fi * F2 + (int64_t)fj * F1;
            for (int cc = 0; cc < 3; ++cc) {
                const int fk = 2 * K + cc;
                if (fk >= nf) continue;
                sum += w[a] * w[bb] * w[cc] * dv(r, row + fk);
            }
        }
    }
    bc[((int64_t)I * nc + J) * nc + K] = static_cast<T>(sum);
}

// Residual and restriction in one pass: bc = R (b - A x) with the values of
// an EpResid launch followed by k_restrict, bit for bit (r rounded to T at
// every fine point, each coarse node's 27 terms added in k_restrict's order),
// without writing r.  A workgroup owns 32 (K) x 16 (J) coarse columns and a
// chunk of `ich` coarse planes and marches over the fine planes 2 Ib .. 2 Ie:
// x of the padded fine region (67 x 35 points) in registers at i-1 / i / i+1
// and in one LDS plane for the j / k neighbours, r of the fine region (65 x 33)
// in LDS; then every thread adds the 9 terms of that plane to its coarse
// node(s): an odd plane 2I+1 is the middle plane of node I, an even plane 2I
// the last plane of node I-1 and the first of node I.
constexpr int kRRThreads = 512, kRRK = 32, kRRJ = 16;
constexpr int kRRFW = 2 * kRRK + 1, kRRFH = 2 * kRRJ + 1;  // fine region of r
constexpr int kRRPW = kRRFW + 2, kRRPH = kRRFH + 2;        // padded region of x
constexpr int kRRSlots = (kRRPW * kRRPH + kRRThreads - 1) / kRRThreads;

template <class T>
__device__ inline void resid_restrict_body(GridPos g, int nf, int nc, int ich, int tiles_k,
                                           int tiles_j, int chunks, double h2inv,
                                           const T* __restrict__ x, const T* __restrict__ b,
                                           T* __restrict__ bc) {
#pragma clang fp contract(off)
    __shared__ T xpl[kRRPH * kRRPW];
    __shared__ T rpl[kRRFH * kRRFW];
    // XCD-aware block -> tile (every 8th block walks consecutive tiles)
    const int G8 = (int)(g.nb >> 3);
    const int lin = (int)(g.bid & 7) * G8 + (int)(g.bid >> 3);
    if (lin >= tiles_k * tiles_j * chunks) return;  // uniform per block
    const int tk = lin % tiles_k, tj = (lin / tiles_k) % tiles_j, ch = lin / (tiles_k * tiles_j);
    const int K0 = tk * kRRK, J0 = tj * kRRJ;
    const int Ib = ch * ich, Ie = min(nc, Ib + ich);
    const int64_t F1 = nf, F2 = F1 * F1;
    const int tid = threadIdx.x;

    // slots of the padded region: byte offset of the fine (j, k) column in a
    // plane (kNoOffset outside the domain), and the r index of interior slots
    // (-1 for the halo ring).  r outside the domain is stored as an exact 0,
    // so the restriction adds its 27 terms without tests: a +0 term leaves
    // the sum (which starts at +0) unchanged.
    const uint32_t plane_bytes = (uint32_t)F2 * (uint32_t)sizeof(T);
    uint32_t off[kRRSlots];
    int ridx[kRRSlots];
#pragma unroll
    for (int s = 0; s < kRRSlots; ++s) {
        const int e = tid + s * kRRThreads;
        const int row = e / kRRPW, col = e % kRRPW;
        const int fj = 2 * J0 - 1 + row, fk = 2 * K0 - 1 + col;
        const bool valid = e < kRRPW * kRRPH;
        off[s] = valid && fj >= 0 && fj < nf && fk >= 0 && fk < nf
                     ? (uint32_t)(fj * nf + fk) * (uint32_t)sizeof(T)
                     : kNoOffset;
        ridx[s] = valid && row >= 1 && row <= kRRFH && col >= 1 && col <= kRRFW
                      ? (row - 1) * kRRFW + (col - 1)
                      : -1;
    }
    // plane q of a (zeros outside the domain)
    auto load_plane = [&](const T* a, int q, T (&v)[kRRSlots]) {
        const bool q_in = (unsigned)q < (unsigned)nf;
        const __amdgpu_buffer_rsrc_t rs =
            plane_rsrc(a + (int64_t)(q_in ? q : 0) * F2, q_in ? plane_bytes : 0u);
#pragma unroll
        for (int s = 0; s < kRRSlots; ++s) v[s] = buf_load<T>(rs, off[s]);
    };

    // coarse node of this thread
    const int cK = K0 + (tid % kRRK), cJ = J0 + tid / kRRK;
    const bool active = cK < nc && cJ < nc;
    const uint32_t coff = active ? (uint32_t)(cJ * nc + cK) * (uint32_t)sizeof(T) : kNoOffset;
    const uint32_t cplane_bytes = (uint32_t)nc * (uint32_t)nc * (uint32_t)sizeof(T);
    const int lr = 2 * (cJ - J0), lc = 2 * (cK - K0);  // its first fine row / column in rpl
    const double w[3] = {0.25, 0.5, 0.25};
    auto add_terms = [&](double& acc, int a) {
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) {
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                acc += w[a] * w[bb] * w[cc] *
                       static_cast<double>(rpl[(lr + bb) * kRRFW + lc + cc]);
        }
    };
    auto store = [&](int I, double acc) {
        buf_store<T>(static_cast<T>(acc), plane_rsrc(bc + (int64_t)I * nc * nc, cplane_bytes),
                     coff);
    };

    const int fb = 2 * Ib, fe = min(2 * Ie, nf - 1);
    T xm[kRRSlots], xc[kRRSlots], xp[kRRSlots], nx[kRRSlots], bv[kRRSlots], nb[kRRSlots];
    load_plane(x, fb - 1, xm);
    load_plane(x, fb, xc);
    load_plane(x, fb + 1, xp);
    load_plane(b, fb, bv);
    double acc = 0.0;
    for (int fi = fb; fi <= fe; ++fi) {
#pragma unroll
        for (int s = 0; s < kRRSlots; ++s) {
            const int e = tid + s * kRRThreads;
            if (e < kRRPW * kRRPH) xpl[e] = xc[s];
        }
        if (fi < fe) {  // next planes in flight during this one
            load_plane(x, fi + 2, nx);
            load_plane(b, fi + 1, nb);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kRRSlots; ++s) {
            if (ridx[s] < 0) continue;
            const int e = tid + s * kRRThreads;
            const double sum =
                lap_num<double>(static_cast<double>(xc[s]), static_cast<double>(xm[s]),
                                static_cast<double>(xp[s]), static_cast<double>(xpl[e - kRRPW]),
                                static_cast<double>(xpl[e + kRRPW]),
                                static_cast<double>(xpl[e - 1]), static_cast<double>(xpl[e + 1]));
            const T r = static_cast<T>(static_cast<double>(bv[s]) - h2inv * sum);
            rpl[ridx[s]] = off[s] != kNoOffset ? r : T(0);
        }
        __syncthreads();
        if ((fi & 1) == 0) {
            if (fi > fb) {  // last plane of node fi/2 - 1
                add_terms(acc, 2);
                store(fi / 2 - 1, acc);
            }
            acc = 0.0;
            if (fi / 2 < Ie) add_terms(acc, 0);  // first plane of node fi/2
        } else {
            add_terms(acc, 1);
        }
#pragma unroll
        for (int s = 0; s < kRRSlots; ++s) {
            xm[s] = xc[s];
            xc[s] = xp[s];
            xp[s] = nx[s];
            bv[s] = nb[s];
        }
    }
    if (fe & 1) store((fe - 1) / 2, acc);  // nf even: the last node has no plane 2I+2
}
template <class T>
__global__ __launch_bounds__(kRRThreads) void k_resid_restrict(int nf, int nc, int ich,
                                                               int tiles_k, int tiles_j, int chunks,
                                                               double h2inv,
                                                               const T* __restrict__ x,
                                                               const T* __restrict__ b,
                                                               T* __restrict__ bc) {
    resid_restrict_body<T>(hw_grid(), nf, nc, ich, tiles_k, tiles_j, chunks, h2inv, x, b, bc);
}

// 1D trilinear terms of a fine index (poisson_mg.hpp:191-220).
__device__ inline int interp_terms(int f, int nc, int ci[2], double wi[2]) {
    if (f & 1) {
        ci[0] = (f - 1) / 2;
        wi[0] = 1.0;
        return 1;
    }
    const int right = f / 2;
    int cnt = 0;
    if (right >= 1) {
        ci[cnt] = right - 1;
        wi[cnt] = 0.5;
        ++cnt;
    }
    if (right < nc) {
        ci[cnt] = right;
        wi[cnt] = 0.5;
        ++cnt;
    }
    return cnt;
}

// x_f += P x_c (poisson_mg.hpp:222-260).  One thread per cell (I,J,K) that
// covers fine nodes {2I,2I+1} x {2J,2J+1} x {2K,2K+1}: those depend only on the
// coarse nodes {I-1,I}^3.  Per axis an odd fine index takes coarse I with
// weight 1, an even one takes I-1 (if I >= 1) and I (if I < nc) with weight
// 1/2; terms are summed in the reference's order (zero-weight terms add 0).
template <class T>
__device__ inline void prolong_add_body(int bx, int by, int bz, int nf, int nc, const T* xc,
                                        T* xf) {
#pragma clang fp contract(off)
    const int K = bx * 64 + (threadIdx.x & 63);
    const int J = by * 4 + (threadIdx.x >> 6);
    const int I = bz;
    const int ncell = (nf + 1) / 2;
    if (K >= ncell || J >= ncell) return;
    const int64_t C1 = nc, C2 = C1 * C1, F1 = nf, F2 = F1 * F1;
    // Every load of the thread (8 coarse, 8 fine values) is issued before the
    // first use; an index out of range loads element 0 and its value is
    // replaced (a branch per load made each one wait for the previous).
    // coarse block c[a][b][e] at (I-1+a, J-1+b, K-1+e), zero where invalid
    T craw[2][2][2];
    bool cok[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int ci = I - 1 + a, cj = J - 1 + b, ck = K - 1 + e;
                cok[a][b][e] = ci >= 0 && ci < nc && cj >= 0 && cj < nc && ck >= 0 && ck < nc;
                craw[a][b][e] = xc[cok[a][b][e] ? ci * C2 + cj * C1 + ck : 0];
            }
    T fraw[2][2][2];
    bool fok[2][2][2];
    int64_t fid[2][2][2];
#pragma unroll
    for (int pi = 0; pi < 2; ++pi)
#pragma unroll
        for (int pj = 0; pj < 2; ++pj)
#pragma unroll
            for (int pk = 0; pk < 2; ++pk) {
                const int fi = 2 * I + pi, fj = 2 * J + pj, fk = 2 * K + pk;
                fok[pi][pj][pk] = fi < nf && fj < nf && fk < nf;
                fid[pi][pj][pk] = fok[pi][pj][pk] ? (int64_t)fi * F2 + (int64_t)fj * F1 + fk : 0;
                fraw[pi][pj][pk] = xf[fid[pi][pj][pk]];
            }
    double c[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e)
                c[a][b][e] = cok[a][b][e] ? static_cast<double>(craw[a][b][e]) : 0.0;
    // weights: [parity][slot] for slot 0 = coarse I-1, 1 = coarse I
    auto wts = [nc](int idx, double w[2]) {
        // even fine index 2*idx
        w[0] = idx >= 1 ? 0.5 : 0.0;
        w[1] = idx < nc ? 0.5 : 0.0;
    };
    double we_i[2], we_j[2], we_k[2];
    wts(I, we_i);
    wts(J, we_j);
    wts(K, we_k);
    const double wo[2] = {0.0, 1.0};
#pragma unroll
    for (int pi = 0; pi < 2; ++pi) {
        const double* wi = pi ? wo : we_i;
#pragma unroll
        for (int pj = 0; pj < 2; ++pj) {
            const double* wj = pj ? wo : we_j;
#pragma unroll
            for (int pk = 0; pk < 2; ++pk) {
                if (!fok[pi][pj][pk]) continue;
                const double* wk = pk ? wo : we_k;
                double sum = 0.0;
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
#pragma unroll
                        for (int e = 0; e < 2; ++e)
                            if (wi[a] != 0.0 && wj[b] != 0.0 && wk[e] != 0.0)
                                sum += wi[a] * wj[b] * wk[e] * c[a][b][e];
                xf[fid[pi][pj][pk]] =
                    static_cast<T>(static_cast<double>(fraw[pi][pj][pk]) + sum);
            }
        }
    }
}
template <class T>
__global__ __launch_bounds__(256) void k_prolong_add(int nf, int nc, const T* xc, T* xf) {
    prolong_add_body<T>(blockIdx.x, blockIdx.y, blockIdx.z, nf, nc, xc, xf);
}

// Coarse solve by one workgroup with the level in LDS (x starts at zero).
//   mode 0: Chebyshev semi-iteration, `iters` steps over [lo, hi] of D^-1 A
//           (iterates double-buffered in LDS: one barrier per step).
//   mode 1: `iters` RB-GS sweeps with alternating order (poisson_mg.hpp:270-274).
//   n == 1: direct x = b/(6 h2inv) (poisson_mg.hpp:267-269).
// n <= 16: thread t owns the column (j, k) = t mod 256 at the planes
// i = 4 (t / 256) + p, p < 4.  Its own values and the i-neighbours inside the
// column live in registers (the same doubles as the LDS copy); the other
// neighbours come from LDS.
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseMaxPts = 4;  // points per thread: n^3 <= 4096
constexpr int kCoarseMaxN3 = kCoarseThreads * kCoarseMaxPts;

template <class T>
__global__ __launch_bounds__(kCoarseThreads) void k_coarse_lds(int n, double h2inv, int mode,
                                                                int iters, double lo, double hi,
                                                                const T* b, T* xout) {
#pragma clang fp contract(off)
    __shared__ double xs[2][kCoarseMaxN3];
    const int N1 = n, N2 = n * n, N = N2 * n;
    if (n == 1) {
        if (threadIdx.x == 0) xout[0] = static_cast<T>(dv(b, 0) / (6.0 * h2inv));
        return;
    }
    const int jk = threadIdx.x & 255, i0 = 4 * (threadIdx.x >> 8);
    const int j = jk / N1, k = jk % N1;
    const bool col = jk < N2;
    double bl[kCoarseMaxPts], dl[kCoarseMaxPts], xr[kCoarseMaxPts], xn[kCoarseMaxPts];
    bool in[kCoarseMaxPts];
    // neighbour masks shared by the column: j-1, j+1, k-1, k+1 inside
    const bool hjm = j > 0, hjp = j + 1 < n, hkm = k > 0, hkp = k + 1 < n;
    // the right-hand side loads are issued before the first use (a point
    // outside the level loads element 0 and is replaced by zero)
    T braw[kCoarseMaxPts];
#pragma unroll
    for (int p = 0; p < kCoarseMaxPts; ++p) {
        const int i = i0 + p;
        in[p] = col && i < n;
        braw[p] = b[in[p] ? i * N2 + jk : 0];
    }
#pragma unroll
    for (int p = 0; p < kCoarseMaxPts; ++p) {
        bl[p] = in[p] ? static_cast<double>(braw[p]) : 0.0;
        dl[p] = xr[p] = xn[p] = 0.0;
    }
    for (int id = threadIdx.x; id < N; id += kCoarseThreads) xs[0][id] = 0.0;
    __syncthreads();
    int cur = 0;
    if (mode == 0) {
        const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
        const double sigma = theta / delta;
        const double dinv = 1.0 / (6.0 * h2inv);
        double rho = 1.0 / sigma;
        for (int it = 0; it < iters; ++it) {
            double c1 = 0.0, c2 = dinv / theta;
            if (it > 0) {
                const double rho_new = 1.0 / (2.0 * sigma - rho);
                c1 = rho_new * rho;
                c2 = 2.0 * rho_new / delta * dinv;
                rho = rho_new;
            }
            const double* xo = xs[cur];
            double* xw = xs[cur ^ 1];
#pragma unroll
            for (int p = 0; p < kCoarseMaxPts; ++p) {
                if (!in[p]) continue;
                const int i = i0 + p, id = i * N2 + jk;
                const double c = xr[p];
                const double im = p > 0 ? xr[p - 1] : (i > 0 ? xo[id - N2] : 0.0);
                const double ip =
                    i + 1 < n ? (p + 1 < kCoarseMaxPts ? xr[p + 1] : xo[id + N2]) : 0.0;
                const double s = lap_num(c, im, ip, hjm ? xo[id - N1] : 0.0,
                                         hjp ? xo[id + N1] : 0.0, hkm ? xo[id - 1] : 0.0,
                                         hkp ? xo[id + 1] : 0.0);
                const double res = bl[p] - h2inv * s;
                const double dn = it > 0 ? c1 * dl[p] + c2 * res : c2 * res;
                dl[p] = dn;
                xn[p] = c + dn;
                xw[id] = xn[p];
            }
#pragma unroll
            for (int p = 0; p < kCoarseMaxPts; ++p) xr[p] = xn[p];
            __syncthreads();
            cur ^= 1;
        }
    } else {
        const double inv = 1.0 / h2inv;
        double* xo = xs[0];
        for (int sweep = 0; sweep < iters; ++sweep) {
            const int order = sweep % 2;
            for (int pass = 0; pass < 2; ++pass) {
                const int color = order == 0 ? pass : 1 - pass;
#pragma unroll
                for (int p = 0; p < kCoarseMaxPts; ++p) {
                    const int i = i0 + p;
                    if (!in[p] || ((i + j + k) & 1) != color) continue;
                    const int id = i * N2 + jk;
                    double s = bl[p] * inv;
                    if (i > 0) s += p > 0 ? xr[p - 1] : xo[id - N2];
                    if (i + 1 < n) s += p + 1 < kCoarseMaxPts ? xr[p + 1] : xo[id + N2];
                    if (hjm) s += xo[id - N1];
                    if (hjp) s += xo[id + N1];
                    if (hkm) s += xo[id - 1];
                    if (hkp) s += xo[id + 1];
                    xr[p] = s / 6.0;
                    xo[id] = xr[p];
                }
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int p = 0; p < kCoarseMaxPts; ++p)
        if (in[p]) xout[(i0 + p) * N2 + jk] = static_cast<T>(xr[p]);
}

// One fused launch of `steps` polynomial smoothing steps (x == nullptr: the
// iterate starts at zero) with R rows per thread (the region is 64 x 16R).
// Row counts: {2, 4} for FP32, {1, 2} for FP64 storage.
// With norm_part, every workgroup also writes the sum of squares of the x it
// stores there; returns the number of workgroups.
// With ev0 / ev1 (no sum of squares) the launch is bracketed by events that
// the dispatch itself stamps (hipExtLaunchKernelGGL: the kernel's own start
// and end, without the queue's dispatch latency).
template <class T, int S, int R>
inline unsigned launch_fused_sr(Ctx& c, int n, int ichunk, double h2inv, const PolySteps& cs,
                                const T* x, const T* b, T* xo, double* norm_part,
                                hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
    int tk, tj, ch;
    const dim3 g = fused_grid<S, R>(n, ichunk, tk, tj, ch);
    if (ev0 && !norm_part) {
        hipExtLaunchKernelGGL((k_poly_fused<T, S, R, false>), g, dim3(kFusedThreads), 0, c.stream,
                              ev0, ev1, 0, n, ichunk, tk, tj, ch, h2inv, cs, x, b, xo, norm_part);
        NUMA_HIP_LAUNCH_CHECK();
        return g.x;
    }
    if (norm_part) {
        // the variants that hold the FP64 accumulator without spilling
        if constexpr (S <= 3 && R == (sizeof(T) == 4 ? 2 : 1))
            hipLaunchKernelGGL((k_poly_fused<T, S, R, true>), g, dim3(kFusedThreads), 0, c.stream,
                               n, ichunk, tk, tj, ch, h2inv, cs, x, b, xo, norm_part);
        else
            throw Error(NUMA_ERR_ARG, "fused sum of squares: at most 3 steps, smallest rows");
    } else
        hipLaunchKernelGGL((k_poly_fused<T, S, R, false>), g, dim3(kFusedThreads), 0, c.stream, n,
                           ichunk, tk, tj, ch, h2inv, cs, x, b, xo, norm_part);
    NUMA_HIP_LAUNCH_CHECK();
    return g.x;
}

template <class T>
constexpr int fused_rows_small() { return sizeof(T) == 4 ? 2 : 1; }

// Box form of a fused launch: rows = -shape selects the output box
// (i x j x k) of k_poly_box; shapes 1..kBoxShapes.
constexpr int kBoxShapes = 3;
constexpr int box_edge_i(int shape) { return shape == 1 ? 8 : 4; }
constexpr int box_edge_j(int shape) { return shape == 3 ? 4 : 8; }
constexpr int box_edge_k(int) { return 16; }

// LDS of the fused kernel: STEPS levels x 2 planes x (16R+2) x 66 values;
// of the box form: 2 copies of the box grown by STEPS on every side.
template <class T>
constexpr bool fused_fits(int steps, int rows) {
    if (rows < 0) {  // box form: up to kFusedMaxSteps steps (a phase of sweeps)
        const int sh = -rows;
        if (sh > kBoxShapes || steps > kFusedMaxSteps) return false;
        return 2 * (int64_t)(box_edge_i(sh) + 2 * steps) * (box_edge_j(sh) + 2 * steps) *
                   (box_edge_k(sh) + 2 * steps) * (int64_t)sizeof(T) <=
               160 * 1024;
    }
    return (int64_t)steps * 2 * ((kFusedThreads / 64) * rows + 2) * 66 * (int64_t)sizeof(T) <=
           160 * 1024;
}

template <class T, int S, int SH>
inline unsigned launch_box_ss(Ctx& c, int n, double h2inv, const PolySteps& cs, const T* x,
                              const T* b, T* xo, double* norm_part) {
    constexpr int EI = box_edge_i(SH), EJ = box_edge_j(SH), EK = box_edge_k(SH);
    if constexpr (fused_fits<T>(S, -SH)) {
        int tk, tj, total;
        const dim3 g = box_grid<EI, EJ, EK>(n, tk, tj, total);
        if (norm_part)
            hipLaunchKernelGGL((k_poly_box<T, S, EI, EJ, EK, true>), g, dim3(kBoxThreads), 0,
                               c.stream, n, tk, tj, total, h2inv, cs, x, b, xo, norm_part);
        else
            hipLaunchKernelGGL((k_poly_box<T, S, EI, EJ, EK, false>), g, dim3(kBoxThreads), 0,
                               c.stream, n, tk, tj, total, h2inv, cs, x, b, xo, norm_part);
        NUMA_HIP_LAUNCH_CHECK();
        return g.x;
    } else {
        throw Error(NUMA_ERR_ARG, "box smoothing: steps per launch x box exceed the LDS");
    }
}

template <class T, int S>
inline unsigned launch_fused_s(Ctx& c, int n, int rows, int ichunk, double h2inv,
                               const PolySteps& cs, const T* x, const T* b, T* xo,
                               double* norm_part, hipEvent_t ev0 = nullptr,
                               hipEvent_t ev1 = nullptr) {
    constexpr int R0 = fused_rows_small<T>();
    switch (rows) {
        case -1: return launch_box_ss<T, S, 1>(c, n, h2inv, cs, x, b, xo, norm_part);
        case -2: return launch_box_ss<T, S, 2>(c, n, h2inv, cs, x, b, xo, norm_part);
        case -3: return launch_box_ss<T, S, 3>(c, n, h2inv, cs, x, b, xo, norm_part);
        default: break;
    }
    if (rows == R0)
        return launch_fused_sr<T, S, R0>(c, n, ichunk, h2inv, cs, x, b, xo, norm_part, ev0, ev1);
    if constexpr (fused_fits<T>(S, 2 * R0)) {
        if (rows == 2 * R0)
            return launch_fused_sr<T, S, 2 * R0>(c, n, ichunk, h2inv, cs, x, b, xo, norm_part,
                                                 ev0, ev1);
    }
    throw Error(NUMA_ERR_ARG, "unsupported steps x rows per thread for fused smoothing");
}

// One fused launch; with norm_part (kMaxPartials doubles) also the sums of
// squares of the written x per workgroup.  Returns the number of workgroups.
template <class T>
inline unsigned launch_fused(Ctx& c, int n, int steps, int rows, int ichunk, double h2inv,
                             const PolySteps& cs, const T* x, const T* b, T* xo,
                             double* norm_part = nullptr, hipEvent_t ev0 = nullptr,
                             hipEvent_t ev1 = nullptr) {
    switch (steps) {
        case 1: return launch_fused_s<T, 1>(c, n, rows, ichunk, h2inv, cs, x, b, xo, norm_part, ev0, ev1);
        case 2: return launch_fused_s<T, 2>(c, n, rows, ichunk, h2inv, cs, x, b, xo, norm_part, ev0, ev1);
        case 3: return launch_fused_s<T, 3>(c, n, rows, ichunk, h2inv, cs, x, b, xo, norm_part, ev0, ev1);
        case 4: return launch_fused_s<T, 4>(c, n, rows, ichunk, h2inv, cs, x, b, xo, norm_part, ev0, ev1);
        case 5: return launch_fused_s<T, 5>(c, n, rows, ichunk, h2inv, cs, x, b, xo, norm_part, ev0, ev1);
        case 6: return launch_fused_s<T, 6>(c, n, rows, ichunk, h2inv, cs, x, b, xo, norm_part, ev0, ev1);
        default: throw Error(NUMA_ERR_ARG, "fused smoothing supports 1..6 steps");
    }
}

// Plane chunk of a fused launch (steps, rows) on an n^3 level that fills the
// device in whole waves of resident workgroups (fill_chunk with the
// kernel's occupancy; the 2*steps halo planes of a chunk are recomputed).
template <class T, int S, int R>
inline int fused_fill_chunk_sr(int n, int cu_count) {
    using G = FusedGeom<S, R>;
    static const int per_cu = blocks_per_cu(k_poly_fused<T, S, R>, kFusedThreads);
    const int64_t tiles = (int64_t)G::tiles_k(n) * G::tiles_j(n);
    return fill_chunk(n, tiles, (int64_t)per_cu * cu_count, 2 * S, 4, 128);
}

template <class T, int S>
inline int fused_fill_chunk_s(int n, int rows, int cu_count) {
    constexpr int R0 = fused_rows_small<T>();
    if constexpr (fused_fits<T>(S, 2 * R0))
        if (rows == 2 * R0) return fused_fill_chunk_sr<T, S, 2 * R0>(n, cu_count);
    return fused_fill_chunk_sr<T, S, R0>(n, cu_count);
}

template <class T>
inline int fused_fill_chunk(int n, int steps, int rows, int cu_count) {
    switch (steps) {
        case 1: return fused_fill_chunk_s<T, 1>(n, rows, cu_count);
        case 2: return fused_fill_chunk_s<T, 2>(n, rows, cu_count);
        case 3: return fused_fill_chunk_s<T, 3>(n, rows, cu_count);
        case 4: return fused_fill_chunk_s<T, 4>(n, rows, cu_count);
        case 5: return fused_fill_chunk_s<T, 5>(n, rows, cu_count);
        default: return fused_fill_chunk_s<T, 6>(n, rows, cu_count);
    }
}

// Exact solve of the coarsest level (n <= 16) by the discrete sine transform:
// the 7-point -Lap_h with zero Dirichlet values is diagonal in the basis
// S_pk = sin(pi (p+1)(k+1)/(n+1)) along each axis, with eigenvalues
// h2inv (c_p + c_q + c_r), c_p = 2 - 2 cos(pi (p+1)/(n+1)), and S S = (n+1)/2 I,
// so x = (2/(n+1))^3 S3 diag(1 / (h2inv (c_p + c_q + c_r))) S3 b.
//
// One workgroup of 16 waves on a 16^3 block zero-padded past n (S is zero
// past n, so padded values stay zero).  A transform along one axis is sixteen
// 16 x 16 x 16 FP64 products, one per wave, each four
// v_mfma_f64_16x16x4f64 steps (lane l: A[l%16][4c + l/16], B[4c + l/16][l%16];
// result D[l/16 + 4r][l%16], r = 0..3); three passes, the scaling, three
// passes back, the block double-buffered in LDS.  `tab` holds S (16 x 16,
// padded) and c (16), computed on the host.
//
// Every pass reads S as S[l%16][4c + l/16] (the k pass as its B operand, the
// j and i passes as A): four values per lane, read from LDS once and held in
// registers.  LDS rows (of the block and of S) are kDstRow = 17 doubles apart:
// the k pass reads its A operand down a column (16 lanes, one row each),
// which at 16 doubles (128 bytes) per row put all 16 lanes on one bank pair
// (about 2 us per pass, 15 us per solve); at 17 they spread over all banks.
// The values and the order of every product are those of the unpadded
// layout.
constexpr int kDstTab = 16 * 16 + 16;
constexpr int kDstRow = 17;

template <class T>
__device__ inline void coarse_dst_body(int n, double h2inv, const double* __restrict__ tab,
                                       const T* __restrict__ b, T* __restrict__ xout) {
#pragma clang fp contract(off)
    __shared__ double u[2][256 * kDstRow];  // row i*16 + j, column k
    __shared__ double S[16][kDstRow];
    __shared__ double cp[16];
    const int t = threadIdx.x, lane = t & 63, g = t >> 6;
    const int r16 = lane & 15, q = lane >> 4;
    // all loads (table and right-hand side) issued before the first LDS write;
    // indices outside the level load element 0 and are replaced
    const double ts = tab[t < 256 ? t : 0];
    const double tc = tab[256 + (t < 16 ? t : 0)];
    T bv[4];
    bool bok[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int id = t + 1024 * m;
        const int i = id >> 8, j = (id >> 4) & 15, k = id & 15;
        bok[m] = i < n && j < n && k < n;
        bv[m] = b[bok[m] ? (i * n + j) * n + k : 0];
    }
    if (t < 256) S[t >> 4][t & 15] = ts;
    if (t < 16) cp[t] = tc;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int id = t + 1024 * m;
        u[0][(id >> 4) * kDstRow + (id & 15)] = bok[m] ? static_cast<double>(bv[m]) : 0.0;
    }
    __syncthreads();
    double sv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) sv[c] = S[r16][4 * c + q];
    using d4 = __attribute__((ext_vector_type(4))) double;
    int cur = 0;
    // Pass with step c's A and B operands from `ld_a(in, c)` / `ld_b(in, c)`
    // and D(row, col) written through `st`; wave g takes product g.
    auto pass = [&](auto ld_a, auto ld_b, auto st) {
        const double* in = u[cur];
        double* out = u[cur ^ 1];
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 4; ++c)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ld_a(in, c), ld_b(in, c), acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) st(out, q + 4 * r, r16, acc[r]);
        __syncthreads();
        cur ^= 1;
    };
    auto s_op = [&](const double*, int c) { return sv[c]; };
    // along k: D[j][p] = sum_k U[g][j][k] S[p][k]
    auto pass_k = [&]() {
        pass([&](const double* in, int c) { return in[(g * 16 + r16) * kDstRow + 4 * c + q]; }, s_op,
             [&](double* out, int row, int col, double v) {
                 out[(g * 16 + row) * kDstRow + col] = v;
             });
    };
    // along j: D[p][k] = sum_j S[p][j] U[g][j][k]
    auto pass_j = [&]() {
        pass(s_op,
             [&](const double* in, int c) { return in[(g * 16 + 4 * c + q) * kDstRow + r16]; },
             [&](double* out, int row, int col, double v) {
                 out[(g * 16 + row) * kDstRow + col] = v;
             });
    };
    // along i: D[p][k] = sum_i S[p][i] U[i][g][k]
    auto pass_i = [&]() {
        pass(s_op,
             [&](const double* in, int c) { return in[((4 * c + q) * 16 + g) * kDstRow + r16]; },
             [&](double* out, int row, int col, double v) {
                 out[(row * 16 + g) * kDstRow + col] = v;
             });
    };
    pass_k();
    pass_j();
    pass_i();
    const double f = 8.0 / ((double)(n + 1) * (double)(n + 1) * (double)(n + 1));
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int id = t + 1024 * m;
        const int i = id >> 8, j = (id >> 4) & 15, k = id & 15;
        double& e = u[cur][(id >> 4) * kDstRow + k];
        if (i < n && j < n && k < n) e = f * (e / (h2inv * (cp[i] + cp[j] + cp[k])));
    }
    __syncthreads();
    pass_k();
    pass_j();
    pass_i();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int id = t + 1024 * m;
        const int i = id >> 8, j = (id >> 4) & 15, k = id & 15;
        if (i < n && j < n && k < n)
            xout[(i * n + j) * n + k] = static_cast<T>(u[cur][(id >> 4) * kDstRow + k]);
    }
}
template <class T>
__global__ __launch_bounds__(1024) void k_coarse_dst(int n, double h2inv,
                                                     const double* __restrict__ tab,
                                                     const T* __restrict__ b,
                                                     T* __restrict__ xout) {
    coarse_dst_body<T>(n, h2inv, tab, b, xout);
}

// Coarse planes per workgroup of the fused residual / restriction onto an
// nc^3 level: the chunk that fills the device in whole waves of resident
// workgroups (a chunk of ich coarse planes marches 2 ich + 1 fine planes).
// Any chunk gives the same values.
template <class T>
inline int resid_restrict_chunk(int64_t nc, int cu_count) {
    static const int per_cu = blocks_per_cu(k_resid_restrict<T>, kRRThreads);
    const int64_t tiles = ((nc + kRRK - 1) / kRRK) * ((nc + kRRJ - 1) / kRRJ);
    return fill_chunk(nc, tiles, (int64_t)per_cu * cu_count, 1, 1, 64);
}

// The table of k_coarse_dst for an n^3 level: S (padded with zeros) and c.
inline std::vector<double> dst_table(int64_t n) {
    std::vector<double> tab(kDstTab, 0.0);
    const double th = M_PI / (double)(n + 1);
    for (int64_t p = 0; p < n; ++p) {
        for (int64_t k = 0; k < n; ++k) tab[(size_t)(p * 16 + k)] = std::sin(th * (double)((p + 1) * (k + 1)));
        tab[(size_t)(256 + p)] = 2.0 - 2.0 * std::cos(th * (double)(p + 1));
    }
    return tab;
}

inline double cheb_lmax(int64_t n) { return 1.0 + std::cos(M_PI / (double(n) + 1.0)); }
inline double cheb_lmin(int64_t n) { return 1.0 - std::cos(M_PI / (double(n) + 1.0)); }

// Chebyshev step coefficients (c1[k], c2[k]) of a degree-deg polynomial
// smoother over [lo, hi] of D^-1 A on a level with D = 6 h2inv.
inline void cheb_step_coefs(double h2inv, int deg, double lo, double hi, std::vector<double>& c1,
                            std::vector<double>& c2) {
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    const double sigma = theta / delta;
    const double dinv = 1.0 / (6.0 * h2inv);
    double rho = 1.0 / sigma;
    c1.assign(deg, 0.0);
    c2.assign(deg, 0.0);
    c2[0] = dinv / theta;
    for (int k = 1; k < deg; ++k) {
        const double rho_new = 1.0 / (2.0 * sigma - rho);
        c1[k] = rho_new * rho;
        c2[k] = 2.0 * rho_new / delta * dinv;
        rho = rho_new;
    }
}

struct MGBase {
    virtual ~MGBase() = default;
    virtual void apply(const void* r, void* z) = 0;
    // apply(), and ||z||_2 (FP64 sum of the squares of the stored z) to the
    // device slot norm: from the last smoothing launch when that is a fused
    // one, else by a separate reduction.
    virtual void apply_norm(const void* r, void* z, double* norm) = 0;
    virtual int64_t levels() const = 0;
    virtual int64_t level_n(int64_t l) const = 0;
    // One finest-level smoothing step on internal buffers (primitive timing);
    // returns its algorithmic HBM bytes.
    virtual double smooth_step_for_timing() = 0;
    // Algorithmic HBM bytes of one V-cycle.
    virtual double vcycle_bytes() const = 0;
    // Smoother fusion chosen for level l (NUMA_FUSE_*; AUTO until tuned), its
    // plane chunk and rows per thread.
    virtual void fusion(int64_t l, int64_t* mode, int64_t* ichunk, int64_t* rows) const = 0;
    // Finishes set-up that needs eager device work (the autotuner's timing), so
    // that apply() only launches kernels (and can be captured in a graph).
    virtual void prepare() = 0;
    // Coarse planes per workgroup of the fused residual / restriction launch
    // on every level: > 0 that chunk, 0 the separate residual and restriction
    // launches, < 0 the default choice.
    virtual void set_resid_restrict(int64_t ich) = 0;
    // Instrumentation: with events on, every finest-level fused smoothing
    // launch in the plane-marching form that reads the iterate (the
    // steady-state sweep, not the first one from zero nor the one that also
    // sums ||z||) gets a start / stop event pair stamped by its dispatch;
    // sweep_time() returns their average duration and count (the stream is
    // synchronised first) and clears them.  Graph replay is off while they
    // are on.
    virtual void set_sweep_events(bool on) = 0;
    virtual bool sweep_events_on() const = 0;
    virtual void sweep_time(double* avg_ms, int64_t* count) = 0;
};

template <class T>
class MG : public MGBase {
public:
    MG(Ctx& c, int64_t n, const numa_mg_config& cfg) : c_(c), cfg_(cfg) {
        require(n > 0, "multigrid needs a positive grid size");
        require(cfg.pre >= 0 && cfg.post >= 0, "multigrid needs non-negative sweep counts");
        require(cfg.smoother >= NUMA_SMOOTH_RBGS && cfg.smoother <= NUMA_SMOOTH_CHEBYSHEV,
                "unknown smoother");
        require(cfg.smoother != NUMA_SMOOTH_CHEBYSHEV || (cfg.degree >= 1 && cfg.cheb_ratio > 1.0),
                "Chebyshev smoother needs degree >= 1 and cheb_ratio > 1");
        require(n <= 2048, "grid too large", NUMA_ERR_OVERFLOW);
        require(cfg.fuse >= NUMA_FUSE_NONE && cfg.fuse <= NUMA_FUSE_AUTO, "unknown fuse mode");
        require(cfg.ichunk >= 0 && cfg.ichunk <= 4096, "ichunk must be in [0, 4096]");
        require(cfg.rows == 0 || cfg.rows == fused_rows_small<T>() ||
                    cfg.rows == 2 * fused_rows_small<T>() ||
                    (cfg.rows < 0 && cfg.rows >= -kBoxShapes),
                "rows must be 0, the storage type's 1x / 2x rows per thread or a box shape "
                "(-1 .. -3)");
        // Hierarchy (poisson_mg.hpp:44-63, plus even coarsening and limits).
        int64_t m = n;
        double h2inv = (double(n) + 1.0) * (double(n) + 1.0);
        while (true) {
            Level L;
            L.n = m;
            L.h2inv = h2inv;
            levels_.push_back(L);
            if (m <= 1) break;
            if (!cfg.even_coarsening && (m % 2 == 0)) break;
            if (m <= std::max<int64_t>(1, cfg.min_coarse_n)) break;
            if (cfg.max_levels > 0 && (int64_t)levels_.size() >= cfg.max_levels) break;
            m = m / 2;
            h2inv /= 4.0;
        }
        for (size_t l = 0; l < levels_.size(); ++l) {
            Level& L = levels_[l];
            const int64_t N = L.n * L.n * L.n;
            L.x = alloc(N);
            L.x2 = alloc(N);
            if (l > 0) L.b = alloc(N);
            L.r = alloc(N);
            if (cfg.smoother == NUMA_SMOOTH_CHEBYSHEV || l + 1 == levels_.size()) L.d = alloc(N);
            // Chebyshev smoothing coefficients (same recurrence as the coarse kernel).
            if (cfg.smoother == NUMA_SMOOTH_CHEBYSHEV) {
                const double hi = cheb_lmax(L.n), lo = hi / cfg.cheb_ratio;
                cheb_coefs(L, (int)cfg.degree, lo, hi, L.c1, L.c2);
            }
            L.fuse = cfg.smoother == NUMA_SMOOTH_RBGS ? (int)NUMA_FUSE_NONE : (int)cfg.fuse;
            L.ichunk = cfg.ichunk > 0 ? (int)cfg.ichunk : 32;
            L.rows = cfg.rows != 0 ? (int)cfg.rows : fused_rows_small<T>();
            if (L.rows < 0) L.ichunk = 0;  // box form: no plane chunk
            if ((L.fuse == NUMA_FUSE_SWEEP || L.fuse == NUMA_FUSE_PHASE) &&
                cfg.smoother != NUMA_SMOOTH_RBGS) {
                require(fused_fits<T>(fused_group(L), L.rows),
                        "fused smoothing: steps per launch x rows per thread (or box) exceed the "
                        "LDS");
                if (cfg.ichunk == 0 && L.rows > 0)
                    L.ichunk = fused_fill_chunk<T>((int)L.n, fused_group(L), L.rows, c.cu_count);
            }
        }
        require(cfg.coarse_iters >= 0 || levels_.back().n <= 16,
                "the exact coarse solve (coarse_iters < 0) needs a coarsest level of at most 16^3");
        if (cfg.coarse_iters < 0) {
            // on the context's stream: a null-stream copy would break a graph
            // capture that another thread runs on its own context
            dst_host_ = dst_table(levels_.back().n);
            void* p = nullptr;
            NUMA_HIP_CHECK(hipMalloc(&p, dst_host_.size() * sizeof(double)));
            allocs_.push_back(p);
            dst_tab_ = static_cast<double*>(p);
            NUMA_HIP_CHECK(hipMemcpyAsync(dst_tab_, dst_host_.data(),
                                          dst_host_.size() * sizeof(double),
                                          hipMemcpyHostToDevice, c.stream));
        }
        set_resid_restrict(-1);
    }
    void set_resid_restrict(int64_t ich) override {
        require(ich <= 1024, "resid/restrict chunk must be <= 1024");
        for (size_t l = 0; l + 1 < levels_.size(); ++l)
            levels_[l].rr_ich = ich < 0 ? rr_chunk(levels_[l + 1].n) : (int)ich;
    }
    ~MG() override {
        for (auto& e : ev_pairs_) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        for (void* p : allocs_) (void)hipFree(p);
    }
    void set_sweep_events(bool on) override { sweep_events_ = on; }
    bool sweep_events_on() const override { return sweep_events_; }
    void sweep_time(double* avg_ms, int64_t* count) override {
        NUMA_HIP_CHECK(hipStreamSynchronize(c_.stream));
        double sum = 0.0;
        for (size_t q = 0; q < ev_used_; ++q) {
            float ms = 0.0f;
            NUMA_HIP_CHECK(hipEventElapsedTime(&ms, ev_pairs_[q].first, ev_pairs_[q].second));
            sum += ms;
        }
        *count = (int64_t)ev_used_;
        *avg_ms = ev_used_ ? sum / (double)ev_used_ : 0.0;
        ev_used_ = 0;
    }

    int64_t levels() const override { return (int64_t)levels_.size(); }
    int64_t level_n(int64_t l) const override { return levels_[(size_t)l].n; }
    void prepare() override {
        if (!tuned_) autotune();
    }
    void fusion(int64_t l, int64_t* mode, int64_t* ichunk, int64_t* rows) const override {
        *mode = levels_[(size_t)l].fuse;
        *ichunk = levels_[(size_t)l].ichunk;
        *rows = levels_[(size_t)l].rows;
    }

    double smooth_step_for_timing() override {
        if (!tuned_) autotune();
        Level& L = levels_[0];
        const int64_t N = L.n * L.n * L.n;
        T* saved_b = L.b;
        L.b = L.r;
        if (cfg_.smoother != NUMA_SMOOTH_RBGS &&
            (L.fuse == NUMA_FUSE_SWEEP || L.fuse == NUMA_FUSE_PHASE)) {
            // one fused launch of the level's group of steps from a nonzero x
            const int steps = fused_group(L);
            PolySteps cs{};
            for (int q = 0; q < steps; ++q) cs.c2[q] = 0.1 / (6.0 * L.h2inv);
            launch_fused<T>(c_, (int)L.n, steps, L.rows, L.ichunk, L.h2inv, cs, (const T*)L.x,
                            (const T*)L.b, L.x2);
            std::swap(L.x, L.x2);
            L.b = saved_b;
            return (double)N * sizeof(T) * 3.0;  // read x, b; write x
        }
        if (cfg_.smoother == NUMA_SMOOTH_RBGS) {
            rbgs_sweep(L, 0);
            L.b = saved_b;
            return (double)N * sizeof(T) * 4.0;  // each colour: read b, x (half), write x (half)
        }
        if (!L.d) {
            poly_step(L, 0.0, 0.1, false, false, false);
            L.b = saved_b;
            return (double)N * sizeof(T) * 3.0;  // read x, b; write xn
        }
        poly_step(L, 0.5, 0.1, true, true, false);
        L.b = saved_b;
        return (double)N * sizeof(T) * 5.0;  // read x, b, d; write d, xn
    }

    double vcycle_bytes() const override {
        // per level: smoothing steps (first step from zero reads b, writes
        // d, x), residual (x, b -> r), restriction (r -> b_c), prolongation
        // (x_c, x -> x), coarsest solve (b -> x); plus the output copy.  A
        // fused launch reads x and b and writes x (from zero: reads b only).
        const double s = sizeof(T);
        double total = 0.0;
        for (size_t l = 0; l < levels_.size(); ++l) {
            const double N = (double)levels_[l].n * levels_[l].n * levels_[l].n;
            if (l + 1 == levels_.size()) {
                total += 2.0 * N * s;
                break;
            }
            const int per_sweep = cfg_.smoother == NUMA_SMOOTH_CHEBYSHEV ? (int)cfg_.degree : 1;
            const int steps = (int)(cfg_.pre + cfg_.post) * per_sweep;
            const int fz = levels_[l].fuse;
            if (cfg_.smoother != NUMA_SMOOTH_RBGS && (fz == NUMA_FUSE_SWEEP || fz == NUMA_FUSE_PHASE) &&
                per_sweep <= kFusedMaxSteps) {
                const int g = fused_group(levels_[l]);
                const int lp = (int)((cfg_.pre * per_sweep + g - 1) / g);
                const int lq = (int)((cfg_.post * per_sweep + g - 1) / g);
                total += N * s * (3.0 * (lp + lq) - (lp > 0 ? 1.0 : 0.0));
                total += N * s * (levels_[l].rr_ich > 0 ? 2.125 : 3.0 + 1.125);
                total += N * s * 2.125;  // prolongation
                continue;
            }
            double step_bytes = cfg_.smoother == NUMA_SMOOTH_CHEBYSHEV ? 5.0 : 3.0;
            if (cfg_.smoother == NUMA_SMOOTH_RBGS) step_bytes = 4.0;
            total += N * s * (steps * step_bytes - (cfg_.pre > 0 ? step_bytes - 3.0 : 0.0));
            if (levels_[l].rr_ich > 0)
                total += N * s * (2.0 + 0.125);  // fused residual / restriction
            else
                total += N * s * (3.0 + 1.0 + 0.125);  // residual; restriction
            total += N * s * (2.0 + 0.125); // prolongation
        }
        const double N0 = (double)levels_[0].n * levels_[0].n * levels_[0].n;
        // output copy, unless the last post-smoothing launch writes the output
        const bool direct = cfg_.smoother != NUMA_SMOOTH_RBGS && cfg_.post > 0 && levels_.size() > 1;
        return total + (direct ? 0.0 : 2.0 * N0 * s);
    }

    // z = V-cycle(r) with zero initial guess (poisson_mg.hpp:69-80).
    void apply(const void* rv, void* zv) override {
        if (!tuned_) autotune();
        const T* r = static_cast<const T*>(rv);
        T* z = static_cast<T*>(zv);
        Level& F = levels_[0];
        F.b = const_cast<T*>(r);
        // the finest level's last post-smoothing launch writes z directly
        // (not when z is r: that launch still reads b)
        final_dst_ = z != r ? z : nullptr;
        final_done_ = false;
        cycle(0);
        if (!final_done_) launch_convert<T, T>(c_, F.n * F.n * F.n, F.x, z);
        final_dst_ = nullptr;
        F.b = nullptr;
    }
    void apply_norm(const void* rv, void* zv, double* norm) override {
        norm_slot_ = norm;
        norm_done_ = false;
        try {
            apply(rv, zv);
        } catch (...) {
            norm_slot_ = nullptr;
            throw;
        }
        norm_slot_ = nullptr;
        if (!norm_done_) {
            const int64_t N = levels_[0].n * levels_[0].n * levels_[0].n;
            launch_dot<T, T>(c_, N, static_cast<const T*>(zv), static_cast<const T*>(zv),
                             Epilogue{EPI_SQRT, 1.0, nullptr, norm, nullptr});
        }
    }

private:
    struct Level {
        int64_t n = 0;
        double h2inv = 0.0;
        T* x = nullptr;
        T* x2 = nullptr;
        T* b = nullptr;
        T* r = nullptr;
        T* d = nullptr;
        std::vector<double> c1, c2;
        int fuse = NUMA_FUSE_NONE;  // NUMA_FUSE_*
        int ichunk = 32;
        int rows = 4;               // rows per thread of fused launches
        int rr_ich = 0;             // coarse planes per workgroup of k_resid_restrict (0: off)
    };

    double* dst_tab_ = nullptr;      // table of the exact coarse solve (coarse_iters < 0)
    std::vector<double> dst_host_;  // its host copy (source of the asynchronous upload)

    T* alloc(int64_t N) {
        void* p = nullptr;
        NUMA_HIP_CHECK(hipMalloc(&p, (size_t)N * sizeof(T)));
        allocs_.push_back(p);
        return static_cast<T*>(p);
    }

    static void cheb_coefs(const Level& L, int deg, double lo, double hi, std::vector<double>& c1,
                           std::vector<double>& c2) {
        cheb_step_coefs(L.h2inv, deg, lo, hi, c1, c2);
    }
    // One step d <- c1 d + c2 (b - A x), x <- x + d (x double-buffered; with
    // dst the new x goes to dst and the level's buffers are left as they are).
    void poly_step(Level& L, double c1, double c2, bool use_d, bool write_d, bool from_zero,
                   T* dst = nullptr) {
        const int64_t N = L.n * L.n * L.n;
        T* const xo = dst ? dst : L.x2;
        if (from_zero) {
            T* d = write_d ? L.d : (T*)nullptr;
            if (aligned2(L.b, sizeof(T)) && aligned2(d, sizeof(T)) && aligned2(xo, sizeof(T)))
                hipLaunchKernelGGL((k_poly_first<true, T>), dim3(stream_blocks(N)), dim3(kBlock), 0,
                                   c_.stream, N, c2, (const T*)L.b, d, xo);
            else
                hipLaunchKernelGGL((k_poly_first<false, T>), dim3(stream_blocks(N)), dim3(kBlock),
                                   0, c_.stream, N, c2, (const T*)L.b, d, xo);
            NUMA_HIP_LAUNCH_CHECK();
        } else {
            const LdVec<T> ld{L.x};
            if (use_d && write_d)
                launch_stencil<false>(c_, L.n, ld, EpPoly<T, true, true>{L.h2inv, c1, c2, L.b, L.d, xo},
                                      Epilogue{});
            else if (use_d)
                launch_stencil<false>(c_, L.n, ld, EpPoly<T, true, false>{L.h2inv, c1, c2, L.b, L.d, xo},
                                      Epilogue{});
            else if (write_d)
                launch_stencil<false>(c_, L.n, ld, EpPoly<T, false, true>{L.h2inv, c1, c2, L.b, L.d, xo},
                                      Epilogue{});
            else
                launch_stencil<false>(c_, L.n, ld, EpPoly<T, false, false>{L.h2inv, c1, c2, L.b, L.d, xo},
                                      Epilogue{});
        }
        if (!dst) std::swap(L.x, L.x2);
    }

    void rbgs_sweep(Level& L, int order) {
        const int64_t N = L.n * L.n * L.n;
        for (int pass = 0; pass < 2; ++pass) {
            const int color = order == 0 ? pass : 1 - pass;
            hipLaunchKernelGGL((k_rbgs<T>), dim3(stream_blocks(N)), dim3(kBlock), 0, c_.stream,
                               (int)L.n, 1.0 / L.h2inv, color, (const T*)L.b, L.x);
            NUMA_HIP_LAUNCH_CHECK();
        }
    }

    // `sweeps` smoothing sweeps; if from_zero the iterate starts at 0.  With
    // last_dst (polynomial smoothers only) the final launch writes last_dst
    // and final_done_ is set.
    void smooth(Level& L, int sweeps, int order, bool from_zero, T* last_dst = nullptr) {
        const int64_t N = L.n * L.n * L.n;
        if (cfg_.smoother == NUMA_SMOOTH_RBGS) {
            if (from_zero) launch_fill<T>(c_, N, T(0), L.x);
            for (int s = 0; s < sweeps; ++s) rbgs_sweep(L, order);
            return;
        }
        if (sweeps == 0) {
            if (from_zero) launch_fill<T>(c_, N, T(0), L.x);
            return;
        }
        if (L.fuse == NUMA_FUSE_SWEEP || L.fuse == NUMA_FUSE_PHASE) {
            if (smooth_fused(L, sweeps, from_zero, last_dst)) return;
        }
        for (int s = 0; s < sweeps; ++s) {
            const bool last_sweep = s + 1 == sweeps;
            if (cfg_.smoother == NUMA_SMOOTH_WJACOBI) {
                T* dst = last_sweep ? last_dst : nullptr;
                poly_step(L, 0.0, cfg_.omega / (6.0 * L.h2inv), false, false, from_zero && s == 0,
                          dst);
                if (dst) final_done_ = true;
            } else {
                const int deg = (int)cfg_.degree;
                for (int k = 0; k < deg; ++k) {
                    T* dst = last_sweep && k + 1 == deg ? last_dst : nullptr;
                    poly_step(L, L.c1[k], L.c2[k], k > 0, k + 1 < deg, from_zero && s == 0 && k == 0,
                              dst);
                    if (dst) final_done_ = true;
                }
            }
        }
    }

    // Steps per fused launch of a level.
    int fused_group(const Level& L) const {
        const int per = cfg_.smoother == NUMA_SMOOTH_WJACOBI ? 1 : (int)cfg_.degree;
        return L.fuse == NUMA_FUSE_SWEEP ? per : (kFusedMaxSteps / per) * per;
    }

    // Fused form of `sweeps` polynomial sweeps: the step sequence of the
    // sweeps (the first step of each sweep starts a new direction, c1 = 0) in
    // launches of one sweep or of as many whole sweeps as fit in 6 steps.
    // Returns false if a sweep is longer than a fused launch can hold.
    bool smooth_fused(Level& L, int sweeps, bool from_zero, T* last_dst = nullptr) {
        const bool wj = cfg_.smoother == NUMA_SMOOTH_WJACOBI;
        const int per = wj ? 1 : (int)cfg_.degree;
        if (per > kFusedMaxSteps) return false;
        std::vector<double> c1s, c2s;
        for (int s = 0; s < sweeps; ++s)
            for (int k = 0; k < per; ++k) {
                c1s.push_back(wj || k == 0 ? 0.0 : L.c1[(size_t)k]);
                c2s.push_back(wj ? cfg_.omega / (6.0 * L.h2inv) : L.c2[(size_t)k]);
            }
        const int total = (int)c1s.size();
        const int group = fused_group(L);
        for (int st = 0; st < total; st += group) {
            const int cnt = std::min(group, total - st);
            PolySteps cs{};
            for (int q = 0; q < cnt; ++q) {
                cs.c1[q] = c1s[(size_t)(st + q)];
                cs.c2[q] = c2s[(size_t)(st + q)];
            }
            T* dst = st + cnt >= total ? last_dst : nullptr;
            // The V-cycle's output launch also gives ||z|| when asked: the box
            // form, or the march with the smallest rows per thread (any rows
            // give the same x; the accumulator spills with more), for sweeps
            // of up to 3 steps; else a separate reduction follows.  FP32
            // switches from 4 to 2 rows for it (equally fast); FP64 only uses
            // it where the level runs 1 row (its 1-row march is slower than a
            // separate reduction).
            constexpr int R0 = fused_rows_small<T>();
            const bool with_norm = dst && norm_slot_ && cnt <= 3 &&
                                   (L.rows <= R0 || sizeof(T) == 4);
            double* np = with_norm ? c_.red.partials : nullptr;
            int rows = L.rows, ichunk = L.ichunk;
            if (with_norm && rows > R0) {
                rows = R0;
                ichunk = fused_fill_chunk<T>((int)L.n, cnt, R0, c_.cu_count);
            }
            // the plane-marching form only (the roofline kernel of the
            // finest level)
            const bool timed_sweep = sweep_events_ && &L == &levels_[0] &&
                                     !(from_zero && st == 0) && !np && rows > 0;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timed_sweep) {
                if (ev_used_ == ev_pairs_.size()) {
                    hipEvent_t a = nullptr, b = nullptr;
                    NUMA_HIP_CHECK(hipEventCreate(&a));
                    NUMA_HIP_CHECK(hipEventCreate(&b));
                    ev_pairs_.push_back({a, b});
                }
                e0 = ev_pairs_[ev_used_].first;
                e1 = ev_pairs_[ev_used_].second;
                ++ev_used_;
            }
            const unsigned nb = launch_fused<T>(
                c_, (int)L.n, cnt, rows, ichunk, L.h2inv, cs,
                (from_zero && st == 0) ? (const T*)nullptr : (const T*)L.x, (const T*)L.b,
                dst ? dst : L.x2, np, e0, e1);
            if (np) {
                c_.finish(nb, Epilogue{EPI_SQRT, 1.0, nullptr, norm_slot_, nullptr});
                norm_done_ = true;
            }
            if (dst)
                final_done_ = true;
            else
                std::swap(L.x, L.x2);
        }
        return true;
    }

    // NUMA_FUSE_AUTO: per level, time one pre-smoothing phase (from zero, on
    // the level's own buffers with b = 1) for every fusion choice and plane
    // chunk, keep the fastest, and remember it in the tune cache.
    void autotune() {
        tuned_ = true;
        const std::string dev = TuneCache::device_key(c_.device);
        for (size_t l = 0; l + 1 < levels_.size(); ++l) {
            Level& L = levels_[l];
            if (L.fuse != NUMA_FUSE_AUTO) continue;
            std::ostringstream key;
            key << dev << "|mg-smooth-v5|" << (sizeof(T) == 4 ? "f32" : "f64") << "|n=" << L.n
                << "|sm=" << cfg_.smoother << "|deg=" << cfg_.degree << "|pre=" << cfg_.pre;
            std::string val;
            const auto held = TuneCache::get().tuning(key.str());
            if (TuneCache::get().lookup(key.str(), val)) {
                int m = 0, ic = 0, rw = 0;
                constexpr int R0 = fused_rows_small<T>();
                const bool parsed = std::sscanf(val.c_str(), "%d,%d,%d", &m, &ic, &rw) == 3;
                const bool march = ic > 0 && ic <= 4096 && (rw == R0 || rw == 2 * R0);
                const bool box = ic == 0 && rw < 0 && rw >= -kBoxShapes;
                if (parsed && m >= NUMA_FUSE_NONE && m <= NUMA_FUSE_PHASE && (march || box)) {
                    L.fuse = m;
                    L.ichunk = ic;
                    L.rows = rw;
                    if (m == NUMA_FUSE_NONE || fused_fits<T>(fused_group(L), rw)) continue;
                    L.fuse = NUMA_FUSE_AUTO;  // an entry this level cannot run: time again
                }
            }
            const int64_t N = L.n * L.n * L.n;
            T* saved_b = L.b;
            L.b = L.r;
            launch_fill<T>(c_, N, T(1), L.r);
            struct Cand {
                int mode, ichunk, rows;
            };
            constexpr int R0 = fused_rows_small<T>();
            std::vector<Cand> cands{{NUMA_FUSE_NONE, 32, R0}};
            for (int rw : {R0, 2 * R0})
                for (int md : {(int)NUMA_FUSE_SWEEP, (int)NUMA_FUSE_PHASE}) {
                    if (md == NUMA_FUSE_PHASE && cfg_.pre <= 1) continue;
                    L.fuse = md;
                    const int g = fused_group(L);
                    if (!fused_fits<T>(g, rw)) continue;
                    // the chunk that fills the device in whole waves, and powers of two
                    const int fill = fused_fill_chunk<T>((int)L.n, g, rw, c_.cu_count);
                    cands.push_back({md, fill, rw});
                    for (int ic : {16, 32, 64})
                        if (ic != fill) cands.push_back({md, ic, rw});
                }
            for (int md : {(int)NUMA_FUSE_SWEEP, (int)NUMA_FUSE_PHASE})
                for (int sh = 1; sh <= kBoxShapes; ++sh) {
                    if (md == NUMA_FUSE_PHASE && cfg_.pre <= 1) continue;
                    L.fuse = md;
                    if (fused_fits<T>(fused_group(L), -sh)) cands.push_back({md, 0, -sh});
                }
            hipEvent_t e0, e1;
            NUMA_HIP_CHECK(hipEventCreate(&e0));
            NUMA_HIP_CHECK(hipEventCreate(&e1));
            // $NUMA_HIP_TUNE_LOG set: every candidate's time per phase to stderr
            const bool log = std::getenv("NUMA_HIP_TUNE_LOG") != nullptr;
            double best = 1e300;
            Cand pick = cands[0];
            for (const Cand& cd : cands) {
                L.fuse = cd.mode;
                L.ichunk = cd.ichunk;
                L.rows = cd.rows;
                smooth(L, (int)cfg_.pre, 0, true);  // warm-up
                // best of 5 trials of 3 phases (clock and cache noise)
                float ms = 1e30f;
                for (int trial = 0; trial < 5; ++trial) {
                    NUMA_HIP_CHECK(hipEventRecord(e0, c_.stream));
                    for (int rep = 0; rep < 3; ++rep) smooth(L, (int)cfg_.pre, 0, true);
                    NUMA_HIP_CHECK(hipEventRecord(e1, c_.stream));
                    NUMA_HIP_CHECK(hipEventSynchronize(e1));
                    float t = 0.0f;
                    NUMA_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
                    ms = std::min(ms, t);
                }
                if (log)
                    std::fprintf(stderr, "numa_hip tune: n=%lld mode=%d ichunk=%d rows=%d %.2f us\n",
                                 (long long)L.n, cd.mode, cd.ichunk, cd.rows, ms * 1e3 / 3.0);
                // a later candidate must win by more than 2% (timing noise
                // would otherwise flip between equal choices from run to run)
                if (ms < 0.98 * best) {
                    best = ms;
                    pick = cd;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            L.b = saved_b;
            L.fuse = pick.mode;
            L.ichunk = pick.ichunk;
            L.rows = pick.rows;
            std::ostringstream v;
            v << pick.mode << "," << pick.ichunk << "," << pick.rows;
            TuneCache::get().store(key.str(), v.str());
        }
        for (Level& L : levels_)
            if (L.fuse == NUMA_FUSE_AUTO) L.fuse = NUMA_FUSE_NONE;  // coarsest level
    }

    void coarse_solve(Level& L) {
        const int64_t N = L.n * L.n * L.n;
        const bool rb = cfg_.smoother == NUMA_SMOOTH_RBGS;
        if (cfg_.coarse_iters < 0) {  // exact (sine-transform) solve, n <= 16
            hipLaunchKernelGGL((k_coarse_dst<T>), dim3(1), dim3(1024), 0, c_.stream, (int)L.n,
                               L.h2inv, (const double*)dst_tab_, (const T*)L.b, L.x);
            NUMA_HIP_LAUNCH_CHECK();
            return;
        }
        if (N <= kCoarseMaxN3) {
            const double hi = cheb_lmax(L.n), lo = cheb_lmin(L.n);
            hipLaunchKernelGGL((k_coarse_lds<T>), dim3(1), dim3(kCoarseThreads), 0, c_.stream,
                               (int)L.n, L.h2inv, rb ? 1 : 0, (int)cfg_.coarse_iters, lo, hi,
                               (const T*)L.b, L.x);
            NUMA_HIP_LAUNCH_CHECK();
            return;
        }
        // Large coarsest level: the same iterations as separate launches.
        if (rb) {
            launch_fill<T>(c_, N, T(0), L.x);
            for (int s = 0; s < (int)cfg_.coarse_iters; ++s) rbgs_sweep(L, s % 2);
            return;
        }
        std::vector<double> c1, c2;
        cheb_coefs(L, (int)std::max<int64_t>(1, cfg_.coarse_iters), cheb_lmin(L.n), cheb_lmax(L.n),
                   c1, c2);
        const int deg = (int)c1.size();
        for (int k = 0; k < deg; ++k) poly_step(L, c1[k], c2[k], k > 0, k + 1 < deg, k == 0);
    }

    // b_C = R (b_L - A x_L): residual (poisson_mg.hpp:128-149) and restriction
    // (:152-188) in one launch, or the two launches when L.rr_ich is 0.
    void resid_restrict(Level& L, Level& C) {
        if (L.rr_ich <= 0) {
            launch_stencil<false>(c_, L.n, LdVec<T>{L.x}, EpResid<T>{L.h2inv, L.b, L.r},
                                  Epilogue{});
            hipLaunchKernelGGL((k_restrict<T>),
                               dim3((unsigned)((C.n + 63) / 64), (unsigned)((C.n + 3) / 4),
                                    (unsigned)C.n),
                               dim3(256), 0, c_.stream, (int)L.n, (int)C.n, (const T*)L.r, C.b);
            NUMA_HIP_LAUNCH_CHECK();
            return;
        }
        const int tiles_k = (int)((C.n + kRRK - 1) / kRRK), tiles_j = (int)((C.n + kRRJ - 1) / kRRJ);
        const int chunks = (int)((C.n + L.rr_ich - 1) / L.rr_ich);
        const int64_t total = (int64_t)tiles_k * tiles_j * chunks;
        hipLaunchKernelGGL((k_resid_restrict<T>), dim3((unsigned)((total + 7) / 8 * 8)),
                           dim3(kRRThreads), 0, c_.stream, (int)L.n, (int)C.n, L.rr_ich, tiles_k,
                           tiles_j, chunks, L.h2inv, (const T*)L.x, (const T*)L.b, C.b);
        NUMA_HIP_LAUNCH_CHECK();
    }

    // Coarse planes per workgroup of the fused residual / restriction: the
    // chunk that fills the device in whole waves of resident workgroups (a
    // chunk of ich coarse planes marches 2 ich + 1 fine planes).
    int rr_chunk(int64_t nc) const { return resid_restrict_chunk<T>(nc, c_.cu_count); }

    // poisson_mg.hpp:262-294
    void cycle(size_t l) {
        Level& L = levels_[l];
        if (l + 1 == levels_.size()) {
            coarse_solve(L);
            return;
        }
        smooth(L, (int)cfg_.pre, 0, true);
        Level& C = levels_[l + 1];
        resid_restrict(L, C);
        cycle(l + 1);
        {
            const int64_t ncell = (L.n + 1) / 2;
            hipLaunchKernelGGL((k_prolong_add<T>),
                               dim3((unsigned)((ncell + 63) / 64), (unsigned)((ncell + 3) / 4),
                                    (unsigned)ncell),
                               dim3(256), 0, c_.stream, (int)L.n, (int)C.n, (const T*)C.x, L.x);
            NUMA_HIP_LAUNCH_CHECK();
        }
        smooth(L, (int)cfg_.post, 1, false, l == 0 ? final_dst_ : nullptr);
    }

    Ctx& c_;
    numa_mg_config cfg_;
    std::vector<Level> levels_;
    std::vector<void*> allocs_;
    bool tuned_ = false;
    T* final_dst_ = nullptr;  // output of the V-cycle being applied
    bool final_done_ = false;
    double* norm_slot_ = nullptr;  // apply_norm: device slot of ||z||
    bool norm_done_ = false;
    bool sweep_events_ = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pairs_;
    size_t ev_used_ = 0;
};

inline MGBase* make_mg(Ctx& c, int64_t n, const numa_mg_config& cfg, int64_t precision) {
    if (precision == NUMA_PRECISION_MIXED) return new MG<float>(c, n, cfg);
    return new MG<double>(c, n, cfg);
}

}  // namespace numa_hip
