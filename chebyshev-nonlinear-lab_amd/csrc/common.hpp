// common.hpp - shared device/host infrastructure of the gfx950 backend:
// error handling, the context (device + stream + reduction scratch), the
// deterministic "last block reduces" protocol and launch-shape helpers.
#pragma once

#include <algorithm>

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <utility>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/numa_hip.h"

namespace numa_hip {

// ---------------------------------------------------------------------------
// Errors.  Internal code throws; every C entry point converts to a status code
// and a thread-local message.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define NUMA_HIP_CHECK(expr)                                                            \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::numa_hip::Error(e_ == hipErrorOutOfMemory ? NUMA_ERR_NOMEM : NUMA_ERR_HIP, \
                                    std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define NUMA_HIP_LAUNCH_CHECK() NUMA_HIP_CHECK(hipGetLastError())

inline void require(bool ok, const std::string& msg, int code = NUMA_ERR_ARG) {
    if (!ok) throw Error(code, msg);
}

// ---------------------------------------------------------------------------
// Reductions.  A reducing kernel writes one FP64 partial per block (fixed
// thread -> element mapping, fixed in-block order); a one-block finish kernel
// on the same stream sums the partials in a fixed order and applies an
// epilogue to the total.  The result is independent of scheduling (bitwise
// repeatable) and needs no device-scope fence in the streaming kernel (a
// release fence per block would write back the XCD's L2 each time).
// Kernels on one stream run in order, so one partial array serves every
// reduction of a context.
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;            // threads per block of streaming kernels
constexpr int kMaxPartials = 65536;    // max blocks of a reducing launch
constexpr int kFinishThreads = 1024;

struct Reduce {
    double* partials;   // [kMaxPartials]
};

// Position of a block in its launch: the hardware's for a plain launch, the
// instance-local one for a launch over an instance table (batch.hpp).
struct GridPos {
    unsigned bid, nb;
};
__device__ inline GridPos hw_grid() { return GridPos{blockIdx.x, gridDim.x}; }

// Epilogues applied by the finish kernel to the reduced total s.
enum EpiKind : int {
    EPI_STORE = 0,       // out[0] = s * mul
    EPI_SQRT = 1,        // out[0] = sqrt(s)
    EPI_NORM_SCALE = 2,  // out[0] = sqrt(s); out2[0] = sqrt(s) > 0 ? 1/sqrt(s) : 0
};

struct Epilogue {
    int64_t kind;    // EpiKind (8 bytes: the struct has no padding bytes)
    double mul;      // scale for EPI_STORE (read from mul_ptr if not null)
    const double* mul_ptr;
    double* out;
    double* out2;
};

// Calls f(std::integral_constant<int, 0>{}) ... f(std::integral_constant<int, U-1>{}).
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int U, class F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, U>{});
}

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Block sum of v; the total is valid in thread 0.  Fixed combination order.
template <int BLOCK>
__device__ inline double block_sum(double v) {
    __shared__ double ws[BLOCK / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < BLOCK / 64; ++w) t += ws[w];
    }
    return t;
}

// Every thread of a reducing kernel calls this with its private sum.
template <int BLOCK>
__device__ inline void block_partial(double acc, const Reduce& red, unsigned bid) {
    const double t = block_sum<BLOCK>(acc);
    if (threadIdx.x == 0) red.partials[bid] = t;
}

// Sum of partials[0, count) in a fixed order, then the epilogue (one block
// of kFinishThreads threads).
__device__ inline void finish_body(const double* partials, unsigned count, const Epilogue& ep) {
    // four loads in flight per round, added in index order (the same sum as a
    // plain strided loop)
    constexpr unsigned K = kFinishThreads;
    double s = 0.0;
    unsigned i = threadIdx.x;
    for (; i + 3 * K < count; i += 4 * K) {
        const double a0 = partials[i], a1 = partials[i + K], a2 = partials[i + 2 * K],
                     a3 = partials[i + 3 * K];
        s += a0;
        s += a1;
        s += a2;
        s += a3;
    }
    if (i + 2 * K < count) {
        const double a0 = partials[i], a1 = partials[i + K], a2 = partials[i + 2 * K];
        s += a0;
        s += a1;
        s += a2;
    } else if (i + K < count) {
        const double a0 = partials[i], a1 = partials[i + K];
        s += a0;
        s += a1;
    } else if (i < count) {
        s += partials[i];
    }
    s = block_sum<kFinishThreads>(s);
    if (threadIdx.x == 0) {
        switch (ep.kind) {
            case EPI_STORE: ep.out[0] = s * (ep.mul_ptr ? ep.mul_ptr[0] : ep.mul); break;
            case EPI_SQRT: ep.out[0] = sqrt(s); break;
            default: {
                const double r = sqrt(s);
                ep.out[0] = r;
                if (ep.out2) ep.out2[0] = r > 0.0 ? 1.0 / r : 0.0;
            }
        }
    }
}

__global__ void __launch_bounds__(kFinishThreads) k_finish(const double* partials, unsigned count,
                                                           Epilogue ep);

// Plane chunk of an i-marching launch over `tiles` tiles per chunk row, with
// `halo` extra planes per chunk and `slots` blocks resident at once: the
// chunk in [lo, hi] minimising waves * (chunk + halo), the time model of
// blocks that each march their chunk (fewer, longer waves of blocks beat a
// last wave that leaves most of the device idle).
inline int fill_chunk(int64_t n, int64_t tiles, int64_t slots, int halo, int lo, int hi) {
    int best = hi;
    double best_cost = 1e300;
    for (int ich = lo; ich <= hi; ++ich) {
        const int64_t chunks = (n + ich - 1) / ich;
        const int64_t waves = (tiles * chunks + slots - 1) / std::max<int64_t>(1, slots);
        const int64_t planes = std::min<int64_t>(ich, n);
        const double cost = (double)waves * (double)(planes + halo);
        if (cost < best_cost) {
            best_cost = cost;
            best = ich;
        }
    }
    return best;
}

// Resident blocks of `kernel` per CU at `threads` threads per block.
template <class K>
inline int blocks_per_cu(K kernel, int threads) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, threads, 0) != hipSuccess ||
        nb < 1)
        nb = 1;
    return nb;
}

// ---------------------------------------------------------------------------
// Context: one device, one stream, reduction scratch and a pinned host
// staging area for scalars.
// ---------------------------------------------------------------------------
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Reduce red{};
    double* dscal = nullptr;      // device scalar slots
    double* hscal = nullptr;      // pinned host mirror
    static constexpr int kScalars = 64;
    int cu_count = 256;
    int live_handles = 0;  // multigrid / solver handles created on this context

    bool owns = true;      // false: stream and scratch belong to the caller
    int max_partials = kMaxPartials;

    explicit Ctx(int dev);
    // A context over a stream and scratch memory that the caller provides
    // (`partials` holds max_partials doubles): it creates, allocates and
    // frees nothing.
    Ctx(int dev, hipStream_t s, double* partials, int max_partials_, double* scalars, int cus)
        : device(dev), stream(s), dscal(scalars), cu_count(cus), owns(false),
          max_partials(max_partials_) {
        red.partials = partials;
    }
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
    ~Ctx();
    // One-block sum of the first `count` partials and the epilogue.
    void finish(unsigned count, const Epilogue& ep);
    // Reads device scalar slots [first, first+count) to the host (synchronises).
    void read_scalars(const double* dev, double* host, int count);
    void sync() { NUMA_HIP_CHECK(hipStreamSynchronize(stream)); }
};

// ---------------------------------------------------------------------------
// Final-solve safeguard of the outer Newton / Chebyshev loops (an addition to
// methods/krylov_newton.hpp:54-114; off when tight_eta = 0, the reference's
// behaviour).  Below T = max(tight_res, rtol_nl) ||F(x0)|| the forcing term
// is at most tight_eta, and the solve counts as converged only once its last
// accepted step started below T (else one more step is taken).  The final
// iterate then carries a linear error of at most tight_eta * T and a
// nonlinear one of order T^2, whatever path the iterations took.
// ---------------------------------------------------------------------------
inline double tight_threshold(const numa_krylov_params& p, double res0) {
    return std::max(p.tight_res, p.rtol_nl) * res0;
}
inline double tight_cap(const numa_krylov_params& p, double eta, double rnorm, double res0) {
    if (p.tight_eta > 0.0 && rnorm <= tight_threshold(p, res0)) return std::min(eta, p.tight_eta);
    return eta;
}
// last_start: ||F|| where the last accepted step started (ignored when
// iters = 0).
inline bool nl_converged(const numa_krylov_params& p, double rnorm, double res0, int64_t iters,
                         double last_start) {
    if (!(rnorm <= p.rtol_nl * res0)) return false;
    if (!(p.tight_eta > 0.0) || iters == 0) return true;
    return last_start <= tight_threshold(p, res0);
}

// Eisenstat-Walker choice 2 with Kelley's safeguards (krylov_newton.hpp:54-65),
// or a constant forcing term.
inline double forcing_ew(const numa_krylov_params& p, int64_t k, double rnorm, double rprev,
                         double eta_prev, double res0) {
    if (p.forcing == NUMA_FORCING_CONSTANT) return p.eta_init;
    if (k == 0) return p.eta_init;
    double eta = p.gamma * std::pow(rnorm / rprev, p.alpha);
    const double safe = p.gamma * std::pow(eta_prev, p.alpha);
    if (safe > 0.1) eta = std::max(eta, safe);
    eta = std::min(eta, p.eta_max);
    eta = std::max(eta, 0.5 * p.rtol_nl * res0 / rnorm);
    return std::min(eta, p.eta_max);
}
inline double forcing(const numa_krylov_params& p, int64_t k, double rnorm, double rprev,
                      double eta_prev, double res0) {
    return tight_cap(p, forcing_ew(p, k, rnorm, rprev, eta_prev, res0), rnorm, res0);
}

// Launch shape of grid-stride streaming kernels over N elements: each thread
// moves 2 x 2 consecutive elements per iteration (16-byte FP64 / 8-byte FP32
// accesses, contiguous across the wave); up to 4096 blocks (16 per CU).
constexpr int kVec = 4;
inline unsigned stream_blocks(int64_t N) {
    int64_t b = (N + (int64_t)kBlock * kVec - 1) / ((int64_t)kBlock * kVec);
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    return (unsigned)b;
}

}  // namespace numa_hip

// Body guards of every C entry point: exceptions become status codes and a
// thread-local message.
#define NUMA_API_BEGIN try {
#define NUMA_API_END                                          \
    }                                                         \
    catch (const ::numa_hip::Error& e) {                      \
        ::numa_hip::set_last_error(e.what());                 \
        return e.code;                                        \
    }                                                         \
    catch (const std::bad_alloc&) {                           \
        ::numa_hip::set_last_error("host allocation failed"); \
        return NUMA_ERR_NOMEM;                                \
    }                                                         \
    catch (const std::exception& e) {                         \
        ::numa_hip::set_last_error(e.what());                 \
        return NUMA_ERR_HIP;                                  \
    }                                                         \
    return NUMA_OK;

// Opaque C handle of a context.
struct numa_hip_ctx {
    numa_hip::Ctx* impl;
};
