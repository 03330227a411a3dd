// batch.hpp - launches over an instance table: one kernel launch applies the
// same operation to every active instance of a batch of Bratu3D problems of
// different grid size and lambda.
//
// A batched launch concatenates the instances' own launch grids: instance i
// owns blocks [start[i], start[i+1]) and runs exactly the blocks, thread
// mapping and per-block partial sums of its single-instance launch (the
// kernel bodies of blas.hpp / stencil.hpp / mg.hpp take the block position
// as an argument), so every instance's values are bit for bit those of the
// same operation launched for it alone.  The per-instance arguments (device
// pointers and per-level constants) live in a device table, one struct per
// instance; the block ranges travel in the kernel arguments (a uniform
// binary search over at most kBatchMax + 1 offsets), and a device flag per
// instance switches it off (its blocks return at once), so one captured
// graph serves every subset of active instances.  Instance ranges start at
// multiples of 8 blocks, so the XCD-aware block -> tile maps of the bodies
// place an instance's tiles on the XCDs as in its own launch.
#pragma once

#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "blas.hpp"
#include "mg.hpp"

namespace numa_hip {

constexpr int kBatchMax = 256;

struct BMap {
    int count;
    unsigned start[kBatchMax + 1];  // first block of instance i; start[count] = grid size
};

template <class Op>
__global__ __launch_bounds__(Op::kThreads) void k_batch(const typename Op::Args* __restrict__ tab,
                                                        const int* __restrict__ act, BMap map) {
    const unsigned b = blockIdx.x;
    int lo = 0, hi = map.count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (map.start[mid] <= b) lo = mid;
        else hi = mid;
    }
    if (act[lo] == 0) return;  // uniform per block
    Op::run(tab[lo], GridPos{b - map.start[lo], map.start[lo + 1] - map.start[lo]});
}

// w = x - a z with a = *a_ptr (the Newton trial point of a line search; the
// arithmetic of LdStep<double, true, false> with a read on the device, so a
// table of these does not change with the step length).
struct LdStepDev {
    const double* x;
    const double* z;
    const double* a_ptr;
    double a;
    struct Raw {
        double x, z;
    };
    __device__ LdStepDev prep() const {
        LdStepDev r = *this;
        r.a = *a_ptr;
        return r;
    }
    __device__ Raw load(int64_t id) const { return {x[id], z[id]}; }
    __device__ P1 make(const Raw& r) const {
#pragma clang fp contract(off)
        double w = r.x;
        w = w - a * r.z;
        return {w};
    }
};

// ---------------------------------------------------------------------------
// Operations.  Args: the per-instance arguments (no padding bytes: tables are
// cached by content); run(args, position) is the single-instance body.
// ---------------------------------------------------------------------------
// The stencil launch of a grid below 128^3, as launch_stencil makes it:
// tk = 0 the LDS-tiled form with k_stencil's plane partials (n > 32), else
// k_stencil with 64 (32) lanes along k and one node per thread.
template <bool REDUCE, class Load, class Epi>
struct OpStencil {
    static constexpr int kThreads = 256;
    struct Args {
        int32_t n, ichunk, gx, gy, gz, tk;
        Load ldr;
        Epi epi;
        Reduce red;
    };
    __device__ static void run(const Args& a, GridPos g) {
        if (a.tk == 0)
            stencil_lds_body<kLdsRows, REDUCE, true, Load, Epi>(g, a.n, a.ichunk, a.gx, a.gy,
                                                                a.gz, a.ldr, a.epi, a.red);
        else if (a.tk == 64)
            stencil_body<64, REDUCE, false, Load, Epi>(g, a.n, a.ichunk, a.gx, a.gy, a.gz, a.ldr,
                                                       a.epi, a.red);
        else
            stencil_body<32, REDUCE, false, Load, Epi>(g, a.n, a.ichunk, a.gx, a.gy, a.gz, a.ldr,
                                                       a.epi, a.red);
    }
};

struct OpFinish {
    static constexpr int kThreads = kFinishThreads;
    struct Args {
        const double* partials;
        int64_t count;
        Epilogue ep;
    };
    __device__ static void run(const Args& a, GridPos) {
        finish_body(a.partials, (unsigned)a.count, a.ep);
    }
};

template <class TA, class TB>
struct OpDot {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t N;
        const TA* a;
        const TB* b;
        Reduce red;
    };
    __device__ static void run(const Args& a, GridPos g) {
        dot_body<true, TA, TB>(g, a.N, a.a, a.b, a.red);
    }
};

template <bool LAST, class T>
struct OpMgs {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t N;
        T* w;
        const T* Vi;
        const double* hi;
        const double* si;
        const T* Vnext;
        Reduce red;
    };
    __device__ static void run(const Args& a, GridPos g) {
        mgs_body<true, LAST, T>(g, a.N, a.w, a.Vi, a.hi, a.si, a.Vnext, a.red);
    }
};

template <class T>
struct OpFill {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t N;
        T* y;
        double v;
    };
    __device__ static void run(const Args& a, GridPos g) {
        fill_body<true, T>(g, a.N, static_cast<T>(a.v), a.y);
    }
};

template <class TI, class TO>
struct OpConvert {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t N;
        const TI* x;
        TO* y;
    };
    __device__ static void run(const Args& a, GridPos g) {
        convert_body<true, TI, TO>(g, a.N, a.x, a.y);
    }
};

template <class TI, class TO>
struct OpConvertNorm {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t N;
        const TI* x;
        TO* y;
        Reduce red;
    };
    __device__ static void run(const Args& a, GridPos g) {
        convert_norm_body<true, TI, TO>(g, a.N, a.x, a.y, a.red);
    }
};

struct OpXnormExp {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t N;
        const double* x;
        double lambda;
        const double* lam_ptr;
        float* E;
        Reduce red;
    };
    __device__ static void run(const Args& a, GridPos g) {
        xnorm_exp_body<true>(g, a.N, a.x, a.lambda, a.lam_ptr, a.E, a.red);
    }
};

// y (=|+=) sum_{k < *count} coef[k] X_k: k_combine with the coefficients, the
// count and the accumulate flag in device memory (the table holds pointers only).
template <class T, class TY>
struct OpCombineDev {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t N;
        const T* v[kMaxComb];
        const double* coef;
        const int* ctl;  // ctl[0] = count, ctl[1] = accumulate
        TY* y;
    };
    __device__ static void run(const Args& a, GridPos g) {
#pragma clang fp contract(off)
        const int count = a.ctl[0], accumulate = a.ctl[1];
        for_each_pair(
            g, a.N,
            [&](int64_t q0, int64_t q1, bool ok0, bool ok1) {
                double s0a = 0.0, s0b = 0.0, s1a = 0.0, s1b = 0.0;
                if (accumulate) {
                    if (ok0) {
                        const Pair<TY> t = ldp<true>(a.y, q0);
                        s0a = (double)t.a;
                        s0b = (double)t.b;
                    }
                    if (ok1) {
                        const Pair<TY> t = ldp<true>(a.y, q1);
                        s1a = (double)t.a;
                        s1b = (double)t.b;
                    }
                }
                for (int k = 0; k < count; ++k) {
                    const double c = a.coef[k];
                    Pair<T> x0{}, x1{};
                    if (ok0) x0 = ldp<true>(a.v[k], q0);
                    if (ok1) x1 = ldp<true>(a.v[k], q1);
                    s0a += c * (double)x0.a;
                    s0b += c * (double)x0.b;
                    s1a += c * (double)x1.a;
                    s1b += c * (double)x1.b;
                }
                if (ok0) stp<true>(a.y, q0, Pair<TY>{static_cast<TY>(s0a), static_cast<TY>(s0b)});
                if (ok1) stp<true>(a.y, q1, Pair<TY>{static_cast<TY>(s1a), static_cast<TY>(s1b)});
            },
            [&](int64_t i) {
                double s = accumulate ? dv(a.y, i) : 0.0;
                for (int k = 0; k < count; ++k) s += a.coef[k] * dv(a.v[k], i);
                a.y[i] = static_cast<TY>(s);
            });
    }
};

template <class T>
struct OpPolyFirst {
    static constexpr int kThreads = kBlock;
    struct Args {
        int64_t N;
        double c2;
        const T* b;
        T* d;
        T* xn;
    };
    __device__ static void run(const Args& a, GridPos g) {
        poly_first_body<true, T>(g, a.N, a.c2, a.b, a.d, a.xn);
    }
};

// One sweep of STEPS (<= 3) smoothing steps in the box form of k_poly_box
// (x null: the iterate starts at zero).
template <class T, int STEPS, int SHAPE>
struct OpPolyBox {
    static constexpr int kThreads = kBoxThreads;
    static constexpr int EI = box_edge_i(SHAPE), EJ = box_edge_j(SHAPE), EK = box_edge_k(SHAPE);
    struct Args {
        int32_t n, tiles_k, tiles_j, total;
        double h2inv;
        PolySteps cs;
        const T* x;
        const T* b;
        T* xo;
    };
    __device__ static void run(const Args& a, GridPos g) {
        poly_box_body<T, STEPS, EI, EJ, EK, false>(g, a.n, a.tiles_k, a.tiles_j, a.total, a.h2inv,
                                                   a.cs, a.x, a.b, a.xo, nullptr);
    }
};

// One sweep of STEPS smoothing steps in the plane-marching form of
// k_poly_fused (regions of 64 x 16R columns, `ichunk` planes per workgroup).
template <class T, int STEPS, int R>
struct OpPolyFused {
    static constexpr int kThreads = kFusedThreads;
    struct Args {
        int32_t n, ichunk, tiles_k, tiles_j, chunks, pad;
        double h2inv;
        PolySteps cs;
        const T* x;
        const T* b;
        T* xo;
    };
    __device__ static void run(const Args& a, GridPos g) {
        poly_fused_body<T, STEPS, R, false>(g, a.n, a.ichunk, a.tiles_k, a.tiles_j, a.chunks,
                                            a.h2inv, a.cs, a.x, a.b, a.xo, nullptr);
    }
};

template <class T>
struct OpResidRestrict {
    static constexpr int kThreads = kRRThreads;
    struct Args {
        int32_t nf, nc, ich, tiles_k, tiles_j, chunks;
        double h2inv;
        const T* x;
        const T* b;
        T* bc;
    };
    __device__ static void run(const Args& a, GridPos g) {
        resid_restrict_body<T>(g, a.nf, a.nc, a.ich, a.tiles_k, a.tiles_j, a.chunks, a.h2inv, a.x,
                               a.b, a.bc);
    }
};

template <class T>
struct OpProlong {  // k_prolong_add's 3-D grid (gx, gy, ncell), x fastest
    static constexpr int kThreads = 256;
    struct Args {
        int32_t nf, nc, gx, gy;
        const T* xc;
        T* xf;
    };
    __device__ static void run(const Args& a, GridPos g) {
        const int bx = (int)(g.bid % (unsigned)a.gx);
        const int by = (int)((g.bid / (unsigned)a.gx) % (unsigned)a.gy);
        const int bz = (int)(g.bid / (unsigned)(a.gx * a.gy));
        prolong_add_body<T>(bx, by, bz, a.nf, a.nc, a.xc, a.xf);
    }
};

template <class T>
struct OpCoarseDst {
    static constexpr int kThreads = 1024;
    struct Args {
        int64_t n;
        double h2inv;
        const double* tab;
        const T* b;
        T* x;
    };
    __device__ static void run(const Args& a, GridPos) {
        coarse_dst_body<T>((int)a.n, a.h2inv, a.tab, a.b, a.x);
    }
};

// ---------------------------------------------------------------------------
// Device tables of per-instance arguments, cached by content: a table is
// uploaded once (from pinned memory that is never rewritten, so a copy
// captured in a graph stays valid) and reused by every later launch with the
// same arguments.
// ---------------------------------------------------------------------------
class ArgArena {
public:
    explicit ArgArena(hipStream_t s) : stream_(s) {}
    ArgArena(const ArgArena&) = delete;
    ArgArena& operator=(const ArgArena&) = delete;
    ~ArgArena() {
        for (void* p : dev_) (void)hipFree(p);
        for (void* p : host_) (void)hipHostFree(p);
    }
    const void* put(const void* data, size_t bytes) {
        std::string key(static_cast<const char*>(data), bytes);
        auto it = cache_.find(key);
        if (it != cache_.end()) return it->second;
        const size_t need = (bytes + 255) / 256 * 256;
        if (need > left_) {
            const size_t chunk = std::max<size_t>(need, kChunk);
            void* d = nullptr;
            void* h = nullptr;
            NUMA_HIP_CHECK(hipMalloc(&d, chunk));
            dev_.push_back(d);
            NUMA_HIP_CHECK(hipHostMalloc(&h, chunk, hipHostMallocDefault));
            host_.push_back(h);
            dcur_ = static_cast<char*>(d);
            hcur_ = static_cast<char*>(h);
            left_ = chunk;
        }
        std::memcpy(hcur_, data, bytes);
        NUMA_HIP_CHECK(hipMemcpyAsync(dcur_, hcur_, bytes, hipMemcpyHostToDevice, stream_));
        const void* out = dcur_;
        dcur_ += need;
        hcur_ += need;
        left_ -= need;
        bytes_ += need;
        cache_.emplace(std::move(key), out);
        return out;
    }
    size_t bytes() const { return bytes_; }
    size_t tables() const { return cache_.size(); }

private:
    static constexpr size_t kChunk = size_t(4) << 20;
    hipStream_t stream_;
    std::unordered_map<std::string, const void*> cache_;
    std::vector<void*> dev_, host_;
    char* dcur_ = nullptr;
    char* hcur_ = nullptr;
    size_t left_ = 0, bytes_ = 0;
};

}  // namespace numa_hip
