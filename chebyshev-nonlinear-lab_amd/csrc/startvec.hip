// startvec.hip - start-vector files of the C1 campaign and the transforms
// that make them.
//
//   * text format: one value per line with 17 significant digits (the
//     reference writes with std::setprecision(max_digits10),
//     experiments/c1/starts/materialize_coarse_start.cpp:18-43) and reads
//     whitespace-separated finite values up to the end of the file with an
//     exact count (apps/run_campaign.cpp:287-302);
//   * SHA-256 of a file (the frozen manifests' *_sha256 columns), FIPS 180-4;
//   * perturbed starts x + rho ||x|| d / ||d||, d ~ N(0,1) from std::mt19937_64
//     seeded with seed*7919 + 13 (run_campaign.cpp:304-327), with the C++
//     standard library's own engine and normal distribution;
//   * the B1 coarse start: trilinear prolongation n_c -> 2 n_c + 1 of a coarse
//     solution (materialize_coarse_start.cpp:60-167), on the device.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"

namespace numa_hip {

namespace {

// ---- SHA-256 (FIPS 180-4) ------------------------------------------------
// Round constants: the first 32 fractional bits of the cube roots of the
// first 64 primes, initial hash: of the square roots of the first 8 primes,
// both computed exactly with integer roots.
unsigned __int128 iroot(unsigned __int128 v, int k) {
    // roots of values below 2^112 are below 2^40, and (2^40)^3 fits 128 bits
    unsigned __int128 lo = 0, hi = (unsigned __int128)1 << 40;
    while (lo < hi) {  // largest r with r^k <= v
        const unsigned __int128 mid = lo + (hi - lo + 1) / 2;
        unsigned __int128 p = mid;
        for (int i = 1; i < k; ++i) p *= mid;
        if (p <= v)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

struct Sha256Tables {
    uint32_t K[64];
    uint32_t H0[8];
    Sha256Tables() {
        int found = 0;
        for (uint32_t p = 2; found < 64; ++p) {
            bool prime = true;
            for (uint32_t d = 2; d * d <= p; ++d)
                if (p % d == 0) {
                    prime = false;
                    break;
                }
            if (!prime) continue;
            K[found] = (uint32_t)iroot((unsigned __int128)p << 96, 3);
            if (found < 8) H0[found] = (uint32_t)iroot((unsigned __int128)p << 64, 2);
            ++found;
        }
    }
};

const Sha256Tables& sha_tables() {
    static const Sha256Tables t;
    return t;
}

inline uint32_t rotr(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }

class Sha256 {
public:
    Sha256() {
        const auto& t = sha_tables();
        std::memcpy(h_, t.H0, sizeof(h_));
    }
    void update(const unsigned char* p, size_t n) {
        len_ += n;
        while (n > 0) {
            const size_t take = std::min(n, (size_t)64 - fill_);
            std::memcpy(buf_ + fill_, p, take);
            fill_ += take;
            p += take;
            n -= take;
            if (fill_ == 64) {
                block(buf_);
                fill_ = 0;
            }
        }
    }
    void finish(char hex[65]) {
        const uint64_t bits = (uint64_t)len_ * 8;
        const unsigned char one = 0x80, zero = 0;
        update(&one, 1);
        while (fill_ != 56) update(&zero, 1);
        unsigned char l[8];
        for (int i = 0; i < 8; ++i) l[i] = (unsigned char)(bits >> (56 - 8 * i));
        update(l, 8);
        for (int i = 0; i < 8; ++i) std::snprintf(hex + 8 * i, 9, "%08x", h_[i]);
        hex[64] = '\0';
    }

private:
    void block(const unsigned char* b) {
        const auto& t = sha_tables();
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = (uint32_t)b[4 * i] << 24 | (uint32_t)b[4 * i + 1] << 16 |
                   (uint32_t)b[4 * i + 2] << 8 | (uint32_t)b[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h_[0], bb = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6],
                 h = h_[7];
        for (int i = 0; i < 64; ++i) {
            const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            const uint32_t ch = (e & f) ^ (~e & g);
            const uint32_t t1 = h + S1 + ch + t.K[i] + w[i];
            const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            const uint32_t mj = (a & bb) ^ (a & c) ^ (bb & c);
            const uint32_t t2 = S0 + mj;
            h = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = bb;
            bb = a;
            a = t1 + t2;
        }
        h_[0] += a;
        h_[1] += bb;
        h_[2] += c;
        h_[3] += d;
        h_[4] += e;
        h_[5] += f;
        h_[6] += g;
        h_[7] += h;
    }
    uint32_t h_[8];
    unsigned char buf_[64];
    size_t fill_ = 0;
    uint64_t len_ = 0;
};

struct FileCloser {
    FILE* f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

// one value per line, 17 significant digits (std::setprecision(max_digits10)
// in the default float format is printf's %.17g)
void write_values(const char* path, const double* x, int64_t count) {
    FILE* f = std::fopen(path, "w");
    require(f != nullptr, std::string("cannot open ") + path + " for writing");
    FileCloser c{f};
    for (int64_t i = 0; i < count; ++i)
        require(std::fprintf(f, "%.17g\n", x[i]) > 0, std::string("write failed: ") + path);
    c.f = nullptr;
    require(std::fclose(f) == 0, std::string("write failed: ") + path);
}

// whitespace-separated values to the end of the file; every value finite and
// exactly `expected` of them (run_campaign.cpp:287-302)
void read_values(const char* path, int64_t expected, double* x) {
    FILE* f = std::fopen(path, "r");
    require(f != nullptr, std::string("cannot open ") + path);
    FileCloser c{f};
    int64_t count = 0;
    char tok[128];
    while (std::fscanf(f, "%127s", tok) == 1) {
        char* end = nullptr;
        errno = 0;
        const double v = std::strtod(tok, &end);
        require(end != tok && *end == '\0', std::string("not a number in ") + path + ": " + tok);
        require(std::isfinite(v), std::string("non-finite value in ") + path);
        require(count < expected, std::string("more values than expected in ") + path);
        x[count++] = v;
    }
    require(count == expected, std::string("fewer values than expected in ") + path);
}

}  // namespace

// B1 coarse start: xf = P xc, nf = 2 nc + 1 (materialize_coarse_start.cpp:
// 60-167): along each axis an odd fine index takes coarse (f-1)/2 with weight
// 1, an even one f = 2I takes I-1 (if I >= 1) and I (if I < nc) with 1/2; the
// present terms summed in (i, j, k) order.
__global__ __launch_bounds__(256) void k_prolong_set(int nf, int nc, const double* xc, double* xf) {
#pragma clang fp contract(off)
    const int64_t N1 = nf, N = N1 * N1 * N1;
    for (int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x; id < N;
         id += (int64_t)gridDim.x * 256) {
        const int f[3] = {(int)(id / (N1 * N1)), (int)((id / N1) % N1), (int)(id % N1)};
        int ci[3][2], cnt[3];
        double wi[3][2];
        for (int a = 0; a < 3; ++a) {
            cnt[a] = 0;
            if (f[a] & 1) {
                ci[a][0] = (f[a] - 1) / 2;
                wi[a][0] = 1.0;
                cnt[a] = 1;
                continue;
            }
            const int right = f[a] / 2;
            if (right >= 1) {
                ci[a][cnt[a]] = right - 1;
                wi[a][cnt[a]++] = 0.5;
            }
            if (right < nc) {
                ci[a][cnt[a]] = right;
                wi[a][cnt[a]++] = 0.5;
            }
        }
        const int64_t C1 = nc;
        double sum = 0.0;
        for (int a = 0; a < cnt[0]; ++a)
            for (int b = 0; b < cnt[1]; ++b)
                for (int c = 0; c < cnt[2]; ++c)
                    sum += wi[0][a] * wi[1][b] * wi[2][c] *
                           xc[((int64_t)ci[0][a] * C1 + ci[1][b]) * C1 + ci[2][c]];
        xf[id] = sum;
    }
}

}  // namespace numa_hip

using numa_hip::require;

extern "C" {

int numa_hip_start_write(const char* path, const double* x_host, int64_t count) {
    NUMA_API_BEGIN
    require(path && (count == 0 || x_host) && count >= 0, "bad argument");
    numa_hip::write_values(path, x_host, count);
    NUMA_API_END
}

int numa_hip_start_read(const char* path, int64_t expected, double* x_host) {
    NUMA_API_BEGIN
    require(path && expected >= 0 && (expected == 0 || x_host), "bad argument");
    numa_hip::read_values(path, expected, x_host);
    NUMA_API_END
}

int numa_hip_sha256(const void* data, int64_t bytes, char* hex_out) {
    NUMA_API_BEGIN
    require(hex_out && bytes >= 0 && (bytes == 0 || data), "bad argument");
    numa_hip::Sha256 h;
    h.update(static_cast<const unsigned char*>(data), (size_t)bytes);
    h.finish(hex_out);
    NUMA_API_END
}

int numa_hip_sha256_file(const char* path, char* hex_out) {
    NUMA_API_BEGIN
    require(path && hex_out, "bad argument");
    FILE* f = std::fopen(path, "rb");
    require(f != nullptr, std::string("cannot open ") + path);
    numa_hip::FileCloser c{f};
    numa_hip::Sha256 h;
    std::vector<unsigned char> buf(1 << 16);
    size_t got;
    while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) h.update(buf.data(), got);
    require(!std::ferror(f), std::string("read failed: ") + path);
    h.finish(hex_out);
    NUMA_API_END
}

int numa_hip_perturb_host(const double* ref_host, int64_t count, double rho, uint64_t seed,
                 double* out_host) {
    NUMA_API_BEGIN
    require(count >= 0 && (count == 0 || (ref_host && out_host)), "bad argument");
    require(rho > 0.0 && std::isfinite(rho), "rho must be positive");
    std::mt19937_64 gen((unsigned long long)seed * 7919ULL + 13ULL);
    std::normal_distribution<double> normal(0.0, 1.0);
    std::vector<double> d((size_t)count);
    for (double& v : d) v = normal(gen);
    double dn = 0.0, rn = 0.0;
    for (int64_t i = 0; i < count; ++i) {
        dn += d[(size_t)i] * d[(size_t)i];
        rn += ref_host[i] * ref_host[i];
    }
    dn = std::sqrt(dn);
    rn = std::sqrt(rn);
    require(dn > 0.0 && rn > 0.0, "perturbation of a zero vector");
    const double scale = rho * rn / dn;
    for (int64_t i = 0; i < count; ++i) out_host[i] = ref_host[i] + scale * d[(size_t)i];
    NUMA_API_END
}

int numa_hip_prolong3d(numa_hip_ctx* ctx, const double* coarse_dev, int64_t nc,
                       double* fine_dev) {
    NUMA_API_BEGIN
    require(ctx && ctx->impl && coarse_dev && fine_dev, "null argument");
    require(nc >= 1 && nc <= 1023, "coarse size must be in [1, 1023]");
    numa_hip::Ctx& c = *ctx->impl;
    NUMA_HIP_CHECK(hipSetDevice(c.device));
    const int64_t nf = 2 * nc + 1, N = nf * nf * nf;
    const unsigned blocks = (unsigned)std::min<int64_t>((N + 255) / 256, 65536);
    hipLaunchKernelGGL(numa_hip::k_prolong_set, dim3(blocks), dim3(256), 0, c.stream, (int)nf,
                       (int)nc, coarse_dev, fine_dev);
    NUMA_HIP_LAUNCH_CHECK();
    c.sync();
    NUMA_API_END
}

}  // extern "C"
