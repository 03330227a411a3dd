// capi.hip - C entry points: context, device memory, and the kernel-level
// Bratu3D / BLAS-1 operations (include/numa_hip.h).
#include <algorithm>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>

#include "blas.hpp"
#include "stencil.hpp"

namespace numa_hip {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

__global__ void __launch_bounds__(kFinishThreads) k_finish(const double* partials, unsigned count,
                                                           Epilogue ep) {
    finish_body(partials, count, ep);
}

void Ctx::finish(unsigned count, const Epilogue& ep) {
    require(count >= 1 && count <= (unsigned)max_partials, "bad reduction size");
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kFinishThreads), 0, stream,
                       (const double*)red.partials, count, ep);
    NUMA_HIP_LAUNCH_CHECK();
}

Ctx::Ctx(int dev) : device(dev) {
    int count = 0;
    NUMA_HIP_CHECK(hipGetDeviceCount(&count));
    require(dev >= 0 && dev < count, "no such HIP device: " + std::to_string(dev));
    NUMA_HIP_CHECK(hipSetDevice(dev));
    hipDeviceProp_t prop{};
    NUMA_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    cu_count = prop.multiProcessorCount;
    NUMA_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    void* p = nullptr;
    NUMA_HIP_CHECK(hipMalloc(&p, kMaxPartials * sizeof(double)));
    red.partials = static_cast<double*>(p);
    NUMA_HIP_CHECK(hipMalloc(&p, kScalars * sizeof(double)));
    dscal = static_cast<double*>(p);
    NUMA_HIP_CHECK(hipHostMalloc(&p, kScalars * sizeof(double), hipHostMallocDefault));
    hscal = static_cast<double*>(p);
    NUMA_HIP_CHECK(hipStreamSynchronize(stream));
}

Ctx::~Ctx() {
    if (!owns) return;
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (red.partials) (void)hipFree(red.partials);
    if (dscal) (void)hipFree(dscal);
    if (hscal) (void)hipHostFree(hscal);
    if (stream) (void)hipStreamDestroy(stream);
}

void Ctx::read_scalars(const double* dev, double* host, int count) {
    require(count <= kScalars, "too many scalars");
    NUMA_HIP_CHECK(hipMemcpyAsync(hscal, dev, count * sizeof(double), hipMemcpyDeviceToHost, stream));
    NUMA_HIP_CHECK(hipStreamSynchronize(stream));
    std::memcpy(host, hscal, count * sizeof(double));
}

}  // namespace numa_hip

using numa_hip::Ctx;
using numa_hip::Epilogue;
using numa_hip::Error;
using numa_hip::require;


namespace {

Ctx& ctx_of(numa_hip_ctx* c) {
    require(c != nullptr && c->impl != nullptr, "null context");
    NUMA_HIP_CHECK(hipSetDevice(c->impl->device));
    return *c->impl;
}

void check_grid(int64_t n) {
    require(n > 0, "grid size must be positive");
    require(n <= 2048, "grid too large", NUMA_ERR_OVERFLOW);
}

// slots of Ctx::dscal used by the kernel-level entry points
enum { K_XNORM = 0, K_VNORM = 1, K_OUT = 2 };

}  // namespace

namespace {
template <class TV, class TO>
void jvp_fd_impl(Ctx& c, int64_t n, double lambda, const double* x, const double* F0, const TV* v,
                 TO* out) {
    const int64_t N = n * n * n;
    numa_hip::launch_dot<double, double>(
        c, N, x, x, Epilogue{numa_hip::EPI_SQRT, 1.0, nullptr, c.dscal + K_XNORM, nullptr});
    numa_hip::launch_dot<TV, TV>(
        c, N, v, v, Epilogue{numa_hip::EPI_SQRT, 1.0, nullptr, c.dscal + K_VNORM, nullptr});
    const numa_hip::FDStep st{std::sqrt(std::numeric_limits<double>::epsilon()), c.dscal + K_XNORM,
                              c.dscal + K_VNORM};
    const double h2inv = (n + 1.0) * (n + 1.0);
    numa_hip::launch_stencil<false>(
        c, n, numa_hip::LdXpHV<TV>{x, v, st, 0.0},
        numa_hip::EpJvFD<TO, false, false>{h2inv, lambda, st, nullptr, F0, out, nullptr, nullptr,
                                           0.0, 1.0},
        Epilogue{});
}

__global__ __launch_bounds__(numa_hip::kBlock) void k_capi_xnorm_exp(int64_t N, const double* x,
                                                                   double lambda, float* E,
                                                                   numa_hip::Reduce red) {
    numa_hip::xnorm_exp_body<false>(numa_hip::hw_grid(), N, x, lambda, nullptr, E, red);
}

// The mixed solver's J v (EpJvFDe): ||x|| and E = lambda e^x in one pass, ||v||,
// then the rearranged quotient.
void jvp_fd_rearranged_impl(Ctx& c, int64_t n, double lambda, const double* x, const float* v,
                            float* E, float* out) {
    const int64_t N = n * n * n;
    const unsigned nb = numa_hip::stream_blocks(N);
    hipLaunchKernelGGL(k_capi_xnorm_exp, dim3(nb), dim3(numa_hip::kBlock), 0, c.stream, N, x,
                       lambda, E, c.red);
    NUMA_HIP_LAUNCH_CHECK();
    c.finish(nb, Epilogue{numa_hip::EPI_SQRT, 1.0, nullptr, c.dscal + K_XNORM, nullptr});
    numa_hip::launch_dot<float, float>(
        c, N, v, v, Epilogue{numa_hip::EPI_SQRT, 1.0, nullptr, c.dscal + K_VNORM, nullptr});
    const numa_hip::FDStep st{std::sqrt(std::numeric_limits<double>::epsilon()), c.dscal + K_XNORM,
                              c.dscal + K_VNORM};
    const double h2inv = (n + 1.0) * (n + 1.0);
    numa_hip::launch_stencil<false>(
        c, n, numa_hip::LdVec<float>{v},
        numa_hip::EpJvFDe<float, false, false>{h2inv, st, nullptr, E, out, nullptr, nullptr, 0.0,
                                               1.0},
        Epilogue{});
}

// ---- scratch of the no-allocation entry point ----------------------------
// The two norms of numa_hip_bratu_jvp_fd_noalloc and their reduction partials
// live in this file-scope device array: one copy per device, part of the
// library image (no allocation at run time), held until the process ends.
// stream_blocks() launches at most 4096 blocks, so a dot never needs more
// partials than kNoAllocPartials, whatever the grid size.
constexpr int kNoAllocPartials = 4096;
constexpr int kNoAllocSlots = 8;
static __device__ __align__(256) double scratch[kNoAllocPartials + kNoAllocSlots];

struct NoAllocDevice {
    double* ptr = nullptr;  // device address of `scratch` on this device
    int cus = 0;
};
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

// The device's copy of `scratch`, resolved once per device with that device
// current (host code never passes the array's own name to a kernel: on the
// host it is only a stand-in for the symbol).
NoAllocDevice noalloc_device(int dev) {
    require(dev >= 0 && dev < kNoAllocMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    NoAllocDevice& d = g_noalloc[dev];
    if (!d.ptr) {
        void* p = nullptr;
        NUMA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "no device address for the scratch array", NUMA_ERR_HIP);
        int cus = 0;
        NUMA_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        d.cus = cus > 0 ? cus : 1;
        d.ptr = static_cast<double*>(p);
    }
    return d;
}

}  // namespace

extern "C" {

const char* numa_hip_last_error(void) { return numa_hip::g_last_error.c_str(); }

int numa_hip_ctx_create(int device, numa_hip_ctx** out) {
    NUMA_API_BEGIN
    require(out != nullptr, "null argument");
    auto h = std::make_unique<numa_hip_ctx>();
    h->impl = new Ctx(device);
    *out = h.release();
    NUMA_API_END
}

int numa_hip_ctx_destroy(numa_hip_ctx* ctx) {
    NUMA_API_BEGIN
    if (ctx) {
        require(ctx->impl == nullptr || ctx->impl->live_handles == 0,
                "context still has live multigrid / solver handles", NUMA_ERR_STATE);
        delete ctx->impl;
        delete ctx;
    }
    NUMA_API_END
}

int numa_hip_ctx_sync(numa_hip_ctx* ctx) {
    NUMA_API_BEGIN
    ctx_of(ctx).sync();
    NUMA_API_END
}

int numa_hip_malloc(numa_hip_ctx* ctx, size_t bytes, void** out_dev) {
    NUMA_API_BEGIN
    require(out_dev != nullptr, "null argument");
    ctx_of(ctx);
    NUMA_HIP_CHECK(hipMalloc(out_dev, bytes));
    NUMA_API_END
}

int numa_hip_free(numa_hip_ctx* ctx, void* dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    c.sync();
    NUMA_HIP_CHECK(hipFree(dev));
    NUMA_API_END
}

int numa_hip_memcpy_h2d(numa_hip_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    NUMA_HIP_CHECK(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, c.stream));
    c.sync();
    NUMA_API_END
}

int numa_hip_memcpy_d2h(numa_hip_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    NUMA_HIP_CHECK(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    NUMA_API_END
}

int numa_hip_bratu_eval(numa_hip_ctx* ctx, int64_t n, double lambda, const double* u_dev,
                        double* F_dev, double* norm_host) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    check_grid(n);
    require(u_dev && F_dev, "null device pointer");
    const double h2inv = (n + 1.0) * (n + 1.0);
    numa_hip::launch_stencil<true>(
        c, n, numa_hip::LdVec<double>{u_dev}, numa_hip::EpF<false>{h2inv, lambda, F_dev, nullptr},
        Epilogue{numa_hip::EPI_SQRT, 1.0, nullptr, c.dscal + K_OUT, nullptr});
    if (norm_host) c.read_scalars(c.dscal + K_OUT, norm_host, 1);
    NUMA_API_END
}

int numa_hip_bratu_jvp(numa_hip_ctx* ctx, int64_t n, double lambda, const double* u_dev,
                       const double* v_dev, double* out_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    check_grid(n);
    require(u_dev && v_dev && out_dev, "null device pointer");
    const double h2inv = (n + 1.0) * (n + 1.0);
    numa_hip::launch_stencil<false>(
        c, n, numa_hip::LdVec<double>{v_dev},
        numa_hip::EpJvAn<double, false, false>{h2inv, lambda, nullptr, u_dev, out_dev, nullptr,
                                               nullptr, 1.0},
        Epilogue{});
    NUMA_API_END
}

int numa_hip_bratu_d2(numa_hip_ctx* ctx, int64_t n, double lambda, const double* u_dev,
                      const double* v_dev, double* out_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    check_grid(n);
    require(u_dev && v_dev && out_dev, "null device pointer");
    const int64_t N = n * n * n;
    hipLaunchKernelGGL(numa_hip::k_d2_analytic, dim3(numa_hip::stream_blocks(N)),
                       dim3(numa_hip::kBlock), 0, c.stream, N, lambda, (const double*)nullptr,
                       u_dev, v_dev, out_dev);
    NUMA_HIP_LAUNCH_CHECK();
    NUMA_API_END
}

int numa_hip_bratu_jacobi_diag(numa_hip_ctx* ctx, int64_t n, double lambda, const double* u_dev,
                               double* d_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    check_grid(n);
    require(u_dev && d_dev, "null device pointer");
    const int64_t N = n * n * n;
    hipLaunchKernelGGL(numa_hip::k_jacobi_diag, dim3(numa_hip::stream_blocks(N)),
                       dim3(numa_hip::kBlock), 0, c.stream, N, (n + 1.0) * (n + 1.0), lambda,
                       (const double*)nullptr, u_dev, d_dev);
    NUMA_HIP_LAUNCH_CHECK();
    NUMA_API_END
}


int numa_hip_bratu_jvp_fd(numa_hip_ctx* ctx, int64_t n, double lambda, const double* x_dev,
                          const double* F0_dev, const double* v_dev, double* out_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    check_grid(n);
    require(x_dev && F0_dev && v_dev && out_dev, "null device pointer");
    jvp_fd_impl<double, double>(c, n, lambda, x_dev, F0_dev, v_dev, out_dev);
    NUMA_API_END
}

int numa_hip_bratu_jvp_fd_mixed(numa_hip_ctx* ctx, int64_t n, double lambda, const double* x_dev,
                                const double* F0_dev, const float* v_dev, float* out_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    check_grid(n);
    require(x_dev && F0_dev && v_dev && out_dev, "null device pointer");
    jvp_fd_impl<float, float>(c, n, lambda, x_dev, F0_dev, v_dev, out_dev);
    NUMA_API_END
}

int numa_hip_bratu_jvp_fd_rearranged(numa_hip_ctx* ctx, int64_t n, double lambda,
                                     const double* x_dev, const float* v_dev, float* E_dev,
                                     float* out_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    check_grid(n);
    require(x_dev && v_dev && E_dev && out_dev, "null device pointer");
    jvp_fd_rearranged_impl(c, n, lambda, x_dev, v_dev, E_dev, out_dev);
    NUMA_API_END
}

int numa_hip_bratu_jvp_fd_noalloc(int device, void* stream, int64_t n, double lambda,
                                  const double* x_dev, const double* F0_dev, const double* v_dev,
                                  double* out_dev) {
    NUMA_API_BEGIN
    check_grid(n);
    require(x_dev && F0_dev && v_dev && out_dev, "null device pointer");
    int count = 0;
    NUMA_HIP_CHECK(hipGetDeviceCount(&count));
    require(device >= 0 && device < count, "no such HIP device: " + std::to_string(device));
    NUMA_HIP_CHECK(hipSetDevice(device));
    const NoAllocDevice d = noalloc_device(device);
    const int64_t N = n * n * n;
    require(numa_hip::stream_blocks(N) <= (unsigned)kNoAllocPartials,
            "reduction exceeds the no-allocation scratch", NUMA_ERR_OVERFLOW);
    Ctx c(device, static_cast<hipStream_t>(stream), d.ptr, kNoAllocPartials,
          d.ptr + kNoAllocPartials, d.cus);
    jvp_fd_impl<double, double>(c, n, lambda, x_dev, F0_dev, v_dev, out_dev);
    NUMA_API_END
}

int numa_hip_bratu_d2_fd(numa_hip_ctx* ctx, int64_t n, double lambda, const double* x_dev,
                         const double* F0_dev, const double* v_dev, double* out_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    check_grid(n);
    require(x_dev && F0_dev && v_dev && out_dev, "null device pointer");
    const int64_t N = n * n * n;
    numa_hip::launch_dot<double, double>(
        c, N, x_dev, x_dev, Epilogue{numa_hip::EPI_SQRT, 1.0, nullptr, c.dscal + K_XNORM, nullptr});
    numa_hip::launch_dot<double, double>(
        c, N, v_dev, v_dev, Epilogue{numa_hip::EPI_SQRT, 1.0, nullptr, c.dscal + K_VNORM, nullptr});
    const numa_hip::FDStep st{std::pow(std::numeric_limits<double>::epsilon(), 0.25),
                              c.dscal + K_XNORM, c.dscal + K_VNORM};
    const double h2inv = (n + 1.0) * (n + 1.0);
    numa_hip::launch_stencil<false>(c, n, numa_hip::LdPM<double>{x_dev, v_dev, st, 0.0},
                                    numa_hip::EpD2FD{h2inv, lambda, st, F0_dev, out_dev, 0.0},
                                    Epilogue{});
    NUMA_API_END
}

int numa_hip_dot(numa_hip_ctx* ctx, int64_t n, const double* a_dev, const double* b_dev,
                 double* result_host) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    require(n >= 0 && result_host, "bad argument");
    if (n == 0) {
        *result_host = 0.0;
    } else {
        numa_hip::launch_dot<double, double>(
            c, n, a_dev, b_dev, Epilogue{numa_hip::EPI_STORE, 1.0, nullptr, c.dscal + K_OUT, nullptr});
        c.read_scalars(c.dscal + K_OUT, result_host, 1);
    }
    NUMA_API_END
}

int numa_hip_dot_f32(numa_hip_ctx* ctx, int64_t n, const float* a_dev, const float* b_dev,
                     double* result_host) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    require(n >= 0 && result_host, "bad argument");
    if (n == 0) {
        *result_host = 0.0;
    } else {
        numa_hip::launch_dot<float, float>(
            c, n, a_dev, b_dev, Epilogue{numa_hip::EPI_STORE, 1.0, nullptr, c.dscal + K_OUT, nullptr});
        c.read_scalars(c.dscal + K_OUT, result_host, 1);
    }
    NUMA_API_END
}

int numa_hip_scal(numa_hip_ctx* ctx, int64_t n, double a, const double* x_dev, double* out_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    require(n >= 0, "bad argument");
    if (n > 0) {
        require(x_dev && out_dev, "null device pointer");
        numa_hip::launch_scale_to(c, n, a, x_dev, out_dev);
    }
    NUMA_API_END
}

int numa_hip_axpy(numa_hip_ctx* ctx, int64_t n, double a, const double* x_dev, double* y_dev) {
    NUMA_API_BEGIN
    Ctx& c = ctx_of(ctx);
    require(n >= 0, "bad argument");
    if (n > 0) numa_hip::launch_axpy<double, double>(c, n, a, x_dev, y_dev);
    NUMA_API_END
}

}  // extern "C"
