// stencil.hpp - 7-point stencil kernels of Bratu3D on gfx950.
//
// Layout: interior nodes only, id = (i*n + j)*n + k, k fastest (the reference's
// layout, problems/bratu3d.hpp:85), zero Dirichlet values outside.
//
// One generic driver k_stencil: a block is 64 (k) x 4 (j) threads; each thread
// owns one (j,k) column of a slab of `ichunk` planes and marches along i,
// keeping the i-1 / i / i+1 values of the field in registers.  The j+-1 and
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

// -Lap_h numerator in the reference's order: 6c - im - ip - jm - jp - km - kp.
__device__ inline double lap_num(double c, double im, double ip, double jm, double jp, double km,
                                 double kp) {
#pragma clang fp contract(off)
    double s = 6.0 * c;
    s -= im;
    s -= ip;
    s -= jm;
    s -= jp;
    s -= km;
    s -= kp;
    return s;
}

template <int BLOCK, bool REDUCE, class Load, class Epi>
__global__ __launch_bounds__(BLOCK) void k_stencil(int n, int ichunk, Load ldr, Epi epi,
                                                    Reduce red, Epilogue ep) {
    const Load L = ldr.prep();
    const Epi E = epi.prep();
    using P = decltype(L(int64_t(0)));
    const int tk = threadIdx.x % kTK, tj = threadIdx.x / kTK;
    const int k = blockIdx.x * kTK + tk, j = blockIdx.y * kTJ + tj;
    const int i0 = blockIdx.z * ichunk;
    const int i1 = min(n, i0 + ichunk);
    double acc = 0.0;
    if (k < n && j < n && i0 < n) {
        const int64_t N1 = n, N2 = N1 * N1;
        int64_t id = ((int64_t)i0 * N1 + j) * N1 + k;
        P pm = i0 > 0 ? L(id - N2) : P::zero();
        P pc = L(id);
        for (int i = i0; i < i1; ++i, id += N2) {
            const P pp = (i + 1 < n) ? L(id + N2) : P::zero();
            const P jm = j > 0 ? L(id - N1) : P::zero();
            const P jp = (j + 1 < n) ? L(id + N1) : P::zero();
            const P km = k > 0 ? L(id - 1) : P::zero();
            const P kp = (k + 1 < n) ? L(id + 1) : P::zero();
            acc += E(id, pc, pm, pp, jm, jp, km, kp);
            pm = pc;
            pc = pp;
        }
    }
    if constexpr (REDUCE) {
        const double t = block_sum<BLOCK>(acc);
        const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        reduce_finish<BLOCK>(t, red, ep, gridDim.x * gridDim.y * gridDim.z, bid);
    }
}

// Planes per thread such that the grid has at most kMaxPartials blocks.
inline int stencil_ichunk(int64_t n) {
    int ich = kIChunk;
    const int64_t gxy = ((n + kTK - 1) / kTK) * ((n + kTJ - 1) / kTJ);
    while (gxy * ((n + ich - 1) / ich) > kMaxPartials) ich *= 2;
    return ich;
}

template <bool REDUCE, class Load, class Epi>
void launch_stencil(Ctx& c, int64_t n, const Load& ldr, const Epi& epi, const Epilogue& ep) {
    const int ich = stencil_ichunk(n);
    const dim3 grid((unsigned)((n + kTK - 1) / kTK), (unsigned)((n + kTJ - 1) / kTJ),
                    (unsigned)((n + ich - 1) / ich));
    hipLaunchKernelGGL((k_stencil<kTK * kTJ, REDUCE, Load, Epi>), grid, dim3(kTK * kTJ), 0,
                       c.stream, (int)n, ich, ldr, epi, c.red, ep);
    NUMA_HIP_LAUNCH_CHECK();
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
template <class T>
struct LdVec {  // w = v
    const T* v;
    __device__ LdVec prep() const { return *this; }
    __device__ P1 operator()(int64_t id) const { return {ld(v, id)}; }
};

// w = (x - a z) - b u   (Newton trial x - lambda z, krylov_newton.hpp:99;
// Chebyshev step x - z - u/2, :166).  z and u may be null.
template <class TZ>
struct LdStep {
    const double* x;
    const TZ* z;
    double a;
    const double* u;
    double b;
    __device__ LdStep prep() const { return *this; }
    __device__ P1 operator()(int64_t id) const {
#pragma clang fp contract(off)
        double w = x[id];
        if (z) w = w - a * ld(z, id);
        if (u) w = w - b * u[id];
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
    __device__ LdXpHV prep() const {
        LdXpHV r = *this;
        r.h = step.h();
        return r;
    }
    __device__ P1 operator()(int64_t id) const {
#pragma clang fp contract(off)
        return {x[id] + h * ld(v, id)};
    }
};

// (x + h v, x + (-h) v) for the central difference (operator.hpp:126-128).
template <class TV>
struct LdPM {
    const double* x;
    const TV* v;
    FDStep step;
    double h;
    __device__ LdPM prep() const {
        LdPM r = *this;
        r.h = step.h();
        return r;
    }
    __device__ P2 operator()(int64_t id) const {
#pragma clang fp contract(off)
        const double xv = x[id], vv = ld(v, id);
        return {xv + h * vv, xv + (-h) * vv};
    }
};

// ---------------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------------
// F = h2inv*(6w - sum nbrs) - lambda e^w (bratu3d.hpp:21-31); optional write of
// the trial point w itself; returns F^2.
struct EpF {
    double h2inv, lambda;
    double* F;
    double* wout;
    __device__ EpF prep() const { return *this; }
    __device__ double operator()(int64_t id, P1 c, P1 im, P1 ip, P1 jm, P1 jp, P1 km,
                                 P1 kp) const {
#pragma clang fp contract(off)
        double f = h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a);
        f -= lambda * exp(c.a);
        F[id] = f;
        if (wout) wout[id] = c.a;
        return f * f;
    }
};

// Forward-difference J v = (F(x+hv) - F0)/h (operator.hpp:113), multiplied by
// the Arnoldi scale *scale_ptr (1 if null), optionally as a residual b - J v.
// Output type TO.  Returns out*dotv[id] (first MGS dot) or out^2 (residual
// norm) or 0.
template <class TO>
struct EpJvFD {
    double h2inv, lambda;
    FDStep step;
    const double* scale_ptr;
    const double* F0;
    TO* out;
    const TO* dotv;
    const double* b;
    double h, scale;
    __device__ EpJvFD prep() const {
        EpJvFD r = *this;
        r.h = step.h();
        r.scale = scale_ptr ? *scale_ptr : 1.0;
        return r;
    }
    __device__ double operator()(int64_t id, P1 c, P1 im, P1 ip, P1 jm, P1 jp, P1 km,
                                 P1 kp) const {
#pragma clang fp contract(off)
        double val = 0.0;
        if (h != 0.0) {  // ||v|| == 0 gives J v = 0 (operator.hpp:105-106)
            double f1 = h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a);
            f1 -= lambda * exp(c.a);
            val = (f1 - F0[id]) / h;
            val = val * scale;
        }
        if (b) val = b[id] - val;
        const TO o = static_cast<TO>(val);
        out[id] = o;
        if (dotv) return static_cast<double>(o) * ld(dotv, id);
        if (b) return static_cast<double>(o) * static_cast<double>(o);
        return 0.0;
    }
};

// Analytic J v = h2inv*(6v - sum nbrs) - lambda e^x v (bratu3d.hpp:33-43),
// same output options as EpJvFD.
template <class TO>
struct EpJvAn {
    double h2inv, lambda;
    const double* scale_ptr;
    const double* x;
    TO* out;
    const TO* dotv;
    const double* b;
    double scale;
    __device__ EpJvAn prep() const {
        EpJvAn r = *this;
        r.scale = scale_ptr ? *scale_ptr : 1.0;
        return r;
    }
    __device__ double operator()(int64_t id, P1 c, P1 im, P1 ip, P1 jm, P1 jp, P1 km,
                                 P1 kp) const {
#pragma clang fp contract(off)
        double val = h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a);
        val -= lambda * exp(x[id]) * c.a;
        val = val * scale;
        if (b) val = b[id] - val;
        const TO o = static_cast<TO>(val);
        out[id] = o;
        if (dotv) return static_cast<double>(o) * ld(dotv, id);
        if (b) return static_cast<double>(o) * static_cast<double>(o);
        return 0.0;
    }
};

// Central-difference F''[v,v] = (F(x+hv) - 2F0 + F(x-hv))/h^2 (operator.hpp:132-133).
struct EpD2FD {
    double h2inv, lambda;
    FDStep step;
    const double* F0;
    double* out;
    double h;
    __device__ EpD2FD prep() const {
        EpD2FD r = *this;
        r.h = step.h();
        return r;
    }
    __device__ double operator()(int64_t id, P2 c, P2 im, P2 ip, P2 jm, P2 jp, P2 km,
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
        out[id] = (fp - 2.0 * F0[id] + fm) / (h * h);
        return 0.0;
    }
};

// Plain -Lap_h v.
template <class TO>
struct EpLap {
    double h2inv;
    TO* out;
    __device__ EpLap prep() const { return *this; }
    __device__ double operator()(int64_t id, P1 c, P1 im, P1 ip, P1 jm, P1 jp, P1 km,
                                 P1 kp) const {
#pragma clang fp contract(off)
        out[id] = static_cast<TO>(h2inv * lap_num(c.a, im.a, ip.a, jm.a, jp.a, km.a, kp.a));
        return 0.0;
    }
};

}  // namespace numa_hip
