// common.hpp - shared device/host infrastructure of the gfx950 backend:
// error handling, the context (device + stream + reduction scratch), the
// deterministic "last block reduces" protocol and launch-shape helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
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
// Reductions.  A reducing kernel writes one FP64 partial per block, counts
// finished blocks with a global atomic, and the last block to finish sums the
// partials in a fixed order and runs an epilogue on the total.  The result is
// therefore independent of block scheduling (bitwise repeatable).
// Kernels on one stream run one after another, so a single partial array and
// counter serve every reduction of a context.
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;           // threads per block of streaming kernels
constexpr int kMaxPartials = 16384;   // max blocks of a reducing launch

struct Reduce {
    double* partials;   // [kMaxPartials]
    unsigned* counter;  // single counter, left at zero by every reduction
};

// Epilogues applied by the last block to the reduced total s.
enum EpiKind : int {
    EPI_STORE = 0,       // out[0] = s * mul
    EPI_SQRT = 1,        // out[0] = sqrt(s)
    EPI_NORM_SCALE = 2,  // out[0] = sqrt(s); out2[0] = sqrt(s) > 0 ? 1/sqrt(s) : 0
};

struct Epilogue {
    int kind;
    double mul;      // scale for EPI_STORE (read from mul_ptr if not null)
    const double* mul_ptr;
    double* out;
    double* out2;
};

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

// Called by every thread of every block of a reducing launch with the
// block-local total (valid in thread 0).  nblocks = number of blocks.
template <int BLOCK>
__device__ inline void reduce_finish(double block_total, const Reduce& red, const Epilogue& ep,
                                     unsigned nblocks, unsigned bid) {
    __shared__ bool is_last;
    if (threadIdx.x == 0) {
        red.partials[bid] = block_total;
        __threadfence();
        const unsigned prev = atomicAdd(red.counter, 1u);
        is_last = (prev == nblocks - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double s = 0.0;
    for (unsigned i = threadIdx.x; i < nblocks; i += BLOCK) {
        s += __builtin_nontemporal_load(&red.partials[i]);
    }
    s = block_sum<BLOCK>(s);
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
        *red.counter = 0u;
    }
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

    explicit Ctx(int dev);
    ~Ctx();
    // Reads device scalar slots [first, first+count) to the host (synchronises).
    void read_scalars(const double* dev, double* host, int count);
    void sync() { NUMA_HIP_CHECK(hipStreamSynchronize(stream)); }
};

// Launch shape of grid-stride streaming kernels over N points: each thread
// handles 4 consecutive elements per iteration; up to 4096 blocks (16 per CU
// of the 256-CU part), capped for the reduction partials.
constexpr int kVec = 4;
inline unsigned stream_blocks(int64_t N) {
    int64_t b = (N + (int64_t)kBlock * kVec - 1) / ((int64_t)kBlock * kVec);
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    return (unsigned)b;
}

}  // namespace numa_hip

// Opaque C handle of a context.
struct numa_hip_ctx {
    numa_hip::Ctx* impl;
};
