// Film compositors on the GPU: linear RGBA film -> 8-bit RGBA frame.
//
// Restates if_compositor::pixel_of, clamp_compositor::commit and aces_compositor::commit with its
// auto exposure (compositor.cpp:29-98):
//   pixel_of(c)   byte = (uint8)(pow(CLAMP(c_k, 0, 1), 1/2.2f) * 255) for r, g, b;
//                 alpha byte = (uint8)(c_3 * 255)
//   clamp         pixel_of(c.at_least(0).at_most(1))            (all four components clamped)
//   aces          x = rgb * e; ldr = x (2.51 x + 0.03) / (x (2.43 x + 0.59) + 0.14) per component;
//                 pixel_of(ldr, alpha)
//   exposure      e = exp(mean over counted pixels of log(0.299 r + 0.587 g + 0.114 b + 1e-3)),
//                 a pixel counted when any of its four components differs from 0 by more than
//                 1e-5 (vec4 != 0.0f, tensor.h:15-17, 121-131)
//
// The log, exp and pow are this file's own: the argument is reduced by powers of two (log) or
// multiples of ln 2 (exp) and the rest evaluated in double precision by its Taylor series
// (atanh series for the log), then rounded once to float. They give the correctly rounded float
// result except within about 1e-14 (relative) of a rounding boundary, are identical on any
// IEEE-754 machine without FMA contraction, and so let the oracle's CPU compositor produce the
// same bytes. The exposure's sum of logs runs in double over a fixed tree (per 1024-pixel chunk:
// four sequential adds per lane then a pairwise tree over 256 lanes; then the same over the
// chunk sums), so it does not depend on scheduling. Values where the reference converts out of
// range or NaN to unsigned char (undefined behaviour) give 0 for NaN and negatives, 255 at or
// above 255.
//
// The 8-bit gamma byte trunc(pow(x, 1/2.2f) * 255) is a non-decreasing step function of x on
// [0, 1] (pow correctly rounded, the product and truncation monotone), so it is the number of its
// 255 thresholds t_k (the smallest float x whose byte is at least k) at or below x. The thresholds
// are found once per process on the host by bisection over float bit patterns with the same pow;
// the kernel estimates the byte with the hardware log2 / exp2 (error far below one step) and
// corrects it against the thresholds in LDS, giving exactly the byte of the exact pow.
//
// Kernels (HBM-bound; 16 B read per pixel per pass, 4 B written by the compose pass):
//   k_exposure_chunks  one block per 1024-pixel chunk writes the chunk's (sum, count).
//   k_exposure_final   one block sums the chunk results in fixed order, writes the exposure.
//   k_compose          one block per 1024 pixels; each lane four pixels 256 apart (coalesced
//                      loads and stores), tonemap + gamma + 8-bit.

#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/x3dpt.h"
#include "render_common.h"

namespace x3d {
int fail(int code, std::string const &msg);
int require_device();
} // namespace x3d

#define X3D_CHECK(call)                                                                            \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return x3d::fail(X3D_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

namespace {

constexpr uint32_t kComposeWG = 256;
constexpr uint32_t kChunkPixels = 1024; // pixels per block, both kernels
constexpr uint64_t kMaxPixels = 1ull << 31;

// ln 2 = kLn2Hi + kLn2Lo, kLn2Hi with 32 significant bits (k * kLn2Hi is exact for |k| < 2^21).
constexpr double kLn2Hi = 0x1.62e42feep-1;
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;
constexpr float kInvGamma = 1.0f / 2.2f;

#define X3D_HD __host__ __device__ inline

X3D_HD double bits_to_double(uint64_t b) {
    double x;
    memcpy(&x, &b, sizeof x);
    return x;
}

X3D_HD uint64_t double_to_bits(double x) {
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    return b;
}

// Natural log in double: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
// s = (m - 1) / (m + 1), |s| <= 0.1716; the series 2 s (1 + s^2/3 + ... + s^22/23) is cut where
// the next term is below 2^-60 of the sum.
X3D_HD double log_d(double x) {
    if (!(x > 0.0))
        return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
    if (x == __builtin_inf())
        return x;
    uint64_t b = double_to_bits(x);
    int e = static_cast<int>(b >> 52) - 1023;
    if (e == -1023) { // subnormal double: scale into the normal range
        b = double_to_bits(x * 0x1p54);
        e = static_cast<int>(b >> 52) - 1023 - 54;
    }
    double m = bits_to_double((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > kSqrt2) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z + 1.0;
    const double lm = (2.0 * s) * p;
    const double de = static_cast<double>(e);
    return de * kLn2Hi + (de * kLn2Lo + lm);
}

// e^t in double: t = k ln 2 + r, |r| <= ln 2 / 2, e^r by its Taylor polynomial of degree 13
// (next term below 5e-18), scaled by 2^k in two steps.
X3D_HD double exp_d(double t) {
    if (t != t)
        return t;
    if (t > 710.0)
        return __builtin_inf();
    if (t < -746.0)
        return 0.0;
    const double k = floor(t * kInvLn2 + 0.5);
    const double r = (t - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 1.0 / 2.0;
    p = p * r + 1.0;
    p = p * r + 1.0;
    const int ki = static_cast<int>(k);
    const int k1 = ki / 2, k2 = ki - k1;
    const double s1 = bits_to_double(static_cast<uint64_t>(k1 + 1023) << 52);
    const double s2 = bits_to_double(static_cast<uint64_t>(k2 + 1023) << 52);
    return p * s1 * s2;
}

// pow(x, 1/2.2f) for x in [0, 1] (or NaN), rounded to float.
X3D_HD float gamma_encode(float x) {
    if (x != x)
        return x;
    if (x == 0.0f)
        return 0.0f;
    if (x == 1.0f)
        return 1.0f;
    return static_cast<float>(exp_d(static_cast<double>(kInvGamma) * log_d(static_cast<double>(x))));
}

X3D_HD uint32_t to_byte(float v) {
    if (!(v > 0.0f))
        return 0u;
    if (v >= 255.0f)
        return 255u;
    return static_cast<uint32_t>(v);
}

X3D_HD float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

// The gamma byte of if_compositor::pixel_of for one channel value (already CLAMPed or NaN), by
// the thresholds: t[k] (k = 1 .. 255) is the smallest float whose byte is at least k, t[256] is
// +inf. NaN and values below t[1] give 0.
__device__ inline uint32_t gamma_byte(float x, const float *t) {
    if (!(x >= t[1]))
        return 0u;
    const float est = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * kInvGamma) * 255.0f;
    int b = static_cast<int>(est);
    b = b < 1 ? 1 : (b > 255 ? 255 : b);
    while (x >= t[b + 1])
        b++;
    while (x < t[b])
        b--;
    return static_cast<uint32_t>(b);
}

// if_compositor::pixel_of (compositor.cpp:29-39), packed r | g << 8 | b << 16 | a << 24.
__device__ inline uint32_t pixel_of(float r, float g, float b, float a, const float *t) {
    const uint32_t R = gamma_byte(clamp01(r), t);
    const uint32_t G = gamma_byte(clamp01(g), t);
    const uint32_t B = gamma_byte(clamp01(b), t);
    const uint32_t A = to_byte(a * 255.0f);
    return R | (G << 8) | (B << 16) | (A << 24);
}

// The thresholds, once per process: for k = 1 .. 255 the smallest float x in (0, 1] with
// to_byte(gamma_encode(x) * 255) >= k, by bisection over the bit patterns of [0, 1] (ordered as
// integers). Each threshold is checked against the bytes of its 64 neighbours on either side.
std::array<float, 257> const &gamma_thresholds() {
    static std::array<float, 257> t;
    static std::once_flag once;
    std::call_once(once, [] {
        auto byte_of = [](uint32_t u) {
            float x;
            memcpy(&x, &u, sizeof x);
            return to_byte(gamma_encode(x) * 255.0f);
        };
        const uint32_t one = 0x3f800000u;
        t[0] = 0.0f;
        for (uint32_t k = 1; k <= 255; k++) {
            uint32_t lo = 0, hi = one; // byte(lo) < k <= byte(hi)
            while (hi - lo > 1) {
                uint32_t mid = lo + (hi - lo) / 2;
                if (byte_of(mid) >= k)
                    hi = mid;
                else
                    lo = mid;
            }
            for (uint32_t d = 1; d <= 64; d++) {
                if ((hi + d <= one && byte_of(hi + d) < k) || (hi >= d && byte_of(hi - d) >= k))
                    t[k] = NAN; // not a step function near the threshold: rejected at upload
            }
            if (t[k] == t[k]) {
                float x;
                memcpy(&x, &hi, sizeof x);
                t[k] = x;
            }
        }
        t[256] = INFINITY;
    });
    return t;
}

// aces_compositor::aces_tonemap per component (compositor.cpp:68-79).
__device__ inline float aces(float x) {
    const float num = x * (x * 2.51f + 0.03f);
    const float den = x * (x * 2.43f + 0.59f) + 0.14f;
    return num / den;
}

__device__ inline float at_least0_at_most1(float x) {
    const float lo = 0.0f > x ? 0.0f : x;
    return 1.0f < lo ? 1.0f : lo;
}

// The pixel's log luminance term of aces_compositor::exposure (compositor.cpp:81-98), and
// whether the pixel counts (any component farther than 1e-5 from 0).
__device__ inline bool counted(float4 c) {
    constexpr float kEps = 1e-5f;
    return !(fabsf(c.x) <= kEps && fabsf(c.y) <= kEps && fabsf(c.z) <= kEps && fabsf(c.w) <= kEps);
}

__device__ inline float log_luminance(float4 c) {
    const float lum = c.x * 0.299f + c.y * 0.587f + c.z * 0.114f;
    return static_cast<float>(log_d(static_cast<double>(lum + 1e-3f)));
}

struct ExposureCtl {
    uint32_t counted; // pixels in the mean
    float exposure;
};

// Pairwise tree over the block's 256 (sum, count) pairs; result in s_sum[0], s_cnt[0].
__device__ inline void block_tree(double *s_sum, uint32_t *s_cnt, double acc, uint32_t n) {
    const uint32_t t = threadIdx.x;
    s_sum[t] = acc;
    s_cnt[t] = n;
    __syncthreads();
    for (uint32_t s = kComposeWG / 2; s > 0; s >>= 1) {
        if (t < s) {
            s_sum[t] = s_sum[t] + s_sum[t + s];
            s_cnt[t] = s_cnt[t] + s_cnt[t + s];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kComposeWG) void k_exposure_chunks(const float4 *film, uint32_t npix, double *chunk_sum,
                                                               uint32_t *chunk_cnt) {
    __shared__ double s_sum[kComposeWG];
    __shared__ uint32_t s_cnt[kComposeWG];
    const uint32_t t = threadIdx.x, c = blockIdx.x;
    float4 q[kChunkPixels / kComposeWG];
    for (uint32_t j = 0; j < kChunkPixels / kComposeWG; j++) {
        const uint32_t p = c * kChunkPixels + j * kComposeWG + t;
        q[j] = p < npix ? film[p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    double acc = 0.0;
    uint32_t n = 0;
    for (uint32_t j = 0; j < kChunkPixels / kComposeWG; j++) {
        if (counted(q[j])) { // padding beyond npix is all zero: not counted
            acc = acc + static_cast<double>(log_luminance(q[j]));
            n++;
        }
    }
    block_tree(s_sum, s_cnt, acc, n);
    if (t == 0) {
        chunk_sum[c] = s_sum[0];
        chunk_cnt[c] = s_cnt[0];
    }
}

__global__ __launch_bounds__(kComposeWG) void k_exposure_final(const double *chunk_sum, const uint32_t *chunk_cnt,
                                                              uint32_t nchunks, ExposureCtl *ctl) {
    __shared__ double s_sum[kComposeWG];
    __shared__ uint32_t s_cnt[kComposeWG];
    const uint32_t t = threadIdx.x;
    double acc = 0.0;
    uint32_t n = 0;
    for (uint32_t k = t; k < nchunks; k += kComposeWG) {
        acc = acc + chunk_sum[k];
        n += chunk_cnt[k];
    }
    block_tree(s_sum, s_cnt, acc, n);
    if (t == 0) {
        const float mean = static_cast<float>(s_sum[0] / static_cast<double>(s_cnt[0]));
        ctl->exposure = static_cast<float>(exp_d(static_cast<double>(mean)));
        ctl->counted = s_cnt[0];
    }
}

__global__ __launch_bounds__(kComposeWG) void k_compose(const float4 *film, uint32_t npix, int32_t kind,
                                                       const ExposureCtl *ctl, float fixed_exposure,
                                                       const float *thresholds, uint32_t *frame) {
    __shared__ float t[257];
    t[threadIdx.x] = thresholds[threadIdx.x];
    if (threadIdx.x == 0)
        t[256] = thresholds[256];
    const float e = ctl ? ctl->exposure : fixed_exposure;
    float4 c[kChunkPixels / kComposeWG];
    for (uint32_t j = 0; j < kChunkPixels / kComposeWG; j++) {
        const uint32_t p = blockIdx.x * kChunkPixels + j * kComposeWG + threadIdx.x;
        c[j] = p < npix ? film[p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();
    for (uint32_t j = 0; j < kChunkPixels / kComposeWG; j++) {
        const uint32_t p = blockIdx.x * kChunkPixels + j * kComposeWG + threadIdx.x;
        if (p >= npix)
            return;
        uint32_t px;
        if (kind == X3D_COMPOSITOR_ACES)
            px = pixel_of(aces(c[j].x * e), aces(c[j].y * e), aces(c[j].z * e), c[j].w, t);
        else
            px = pixel_of(at_least0_at_most1(c[j].x), at_least0_at_most1(c[j].y), at_least0_at_most1(c[j].z),
                          at_least0_at_most1(c[j].w), t);
        frame[p] = px;
    }
}

} // namespace

struct x3d_compositor {
    int32_t kind = X3D_COMPOSITOR_ACES;
    int32_t auto_exposure = 1; // aces_compositor defaults (compositor.h:51-52)
    float exposure = 0.0f;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    x3d::DevBuf film, frame, chunks, ctl, thresholds;
    ~x3d_compositor() {
        if (ev0)
            (void)hipEventDestroy(ev0);
        if (ev1)
            (void)hipEventDestroy(ev1);
        if (ev2)
            (void)hipEventDestroy(ev2);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

namespace {

int check_size(uint32_t width, uint32_t height) {
    const uint64_t npix = static_cast<uint64_t>(width) * height;
    if (npix == 0)
        return x3d::fail(X3D_ERR_INVALID_ARGUMENT, "empty film");
    if (npix > kMaxPixels)
        return x3d::fail(X3D_ERR_INVALID_ARGUMENT, "film larger than 2^31 pixels");
    return X3D_OK;
}

int commit_impl(x3d_compositor *c, const float *d_rgba, uint32_t npix, uint8_t *d_frame, x3d_compose_stats *stats) {
    auto t0 = std::chrono::steady_clock::now();
    const bool auto_e = c->kind == X3D_COMPOSITOR_ACES && c->auto_exposure;
    const uint32_t nchunks = (npix + kChunkPixels - 1) / kChunkPixels;
    const float4 *film = reinterpret_cast<const float4 *>(d_rgba);
    ExposureCtl *ctl = nullptr;
    X3D_CHECK(hipEventRecord(c->ev0, c->stream));
    if (auto_e) {
        X3D_CHECK(c->ctl.ensure(256));
        X3D_CHECK(c->chunks.ensure(static_cast<size_t>(nchunks) * (sizeof(double) + sizeof(uint32_t))));
        double *chunk_sum = static_cast<double *>(c->chunks.p);
        uint32_t *chunk_cnt = reinterpret_cast<uint32_t *>(chunk_sum + nchunks);
        ctl = static_cast<ExposureCtl *>(c->ctl.p);
        hipLaunchKernelGGL(k_exposure_chunks, dim3(nchunks), dim3(kComposeWG), 0, c->stream, film, npix, chunk_sum,
                           chunk_cnt);
        hipLaunchKernelGGL(k_exposure_final, dim3(1), dim3(kComposeWG), 0, c->stream,
                           static_cast<const double *>(chunk_sum), static_cast<const uint32_t *>(chunk_cnt), nchunks,
                           ctl);
        X3D_CHECK(hipGetLastError());
    }
    X3D_CHECK(hipEventRecord(c->ev1, c->stream));
    hipLaunchKernelGGL(k_compose, dim3(nchunks), dim3(kComposeWG), 0, c->stream, film, npix, c->kind,
                       static_cast<const ExposureCtl *>(ctl), c->exposure, static_cast<const float *>(c->thresholds.p),
                       reinterpret_cast<uint32_t *>(d_frame));
    X3D_CHECK(hipGetLastError());
    X3D_CHECK(hipEventRecord(c->ev2, c->stream));
    X3D_CHECK(hipEventSynchronize(c->ev2));
    if (stats) {
        ExposureCtl h{};
        if (auto_e)
            X3D_CHECK(hipMemcpy(&h, ctl, sizeof h, hipMemcpyDeviceToHost));
        float ms_e = 0.0f, ms_c = 0.0f;
        X3D_CHECK(hipEventElapsedTime(&ms_e, c->ev0, c->ev1));
        X3D_CHECK(hipEventElapsedTime(&ms_c, c->ev1, c->ev2));
        stats->exposure = auto_e ? h.exposure : (c->kind == X3D_COMPOSITOR_ACES ? c->exposure : 1.0f);
        stats->counted = auto_e ? h.counted : 0u;
        stats->exposure_ms = ms_e;
        stats->compose_ms = ms_c;
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return X3D_OK;
}

} // namespace

extern "C" {

int x3d_compositor_create(int32_t kind, x3d_compositor **out) {
    if (!out)
        return x3d::fail(X3D_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    if (kind != X3D_COMPOSITOR_CLAMP && kind != X3D_COMPOSITOR_ACES)
        return x3d::fail(X3D_ERR_INVALID_ARGUMENT, "unknown compositor kind");
    if (int rc = x3d::require_device())
        return rc;
    auto *c = new x3d_compositor();
    c->kind = kind;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&c->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&c->ev1);
    if (e == hipSuccess)
        e = hipEventCreate(&c->ev2);
    std::array<float, 257> const &t = gamma_thresholds();
    if (e == hipSuccess)
        e = c->thresholds.ensure(sizeof t);
    if (e == hipSuccess)
        e = hipMemcpy(c->thresholds.p, t.data(), sizeof t, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete c;
        return x3d::fail(X3D_ERR_DEVICE, std::string("compositor set-up: ") + hipGetErrorString(e));
    }
    for (float x : t)
        if (x != x) {
            delete c;
            return x3d::fail(X3D_ERR_DEVICE, "gamma byte is not a step function near a threshold");
        }
    *out = c;
    return X3D_OK;
}

int x3d_compositor_destroy(x3d_compositor *c) {
    delete c;
    return X3D_OK;
}

int x3d_compositor_enable_auto_exposure(x3d_compositor *c, int32_t on) {
    if (!c)
        return x3d::fail(X3D_ERR_INVALID_ARGUMENT, "compositor is null");
    c->auto_exposure = on ? 1 : 0;
    return X3D_OK;
}

int x3d_compositor_set_exposure(x3d_compositor *c, float exposure) {
    if (!c)
        return x3d::fail(X3D_ERR_INVALID_ARGUMENT, "compositor is null");
    c->exposure = exposure;
    return X3D_OK;
}

int x3d_compositor_commit(x3d_compositor *c, const float *rgba, uint32_t width, uint32_t height, uint8_t *pixels,
                          x3d_compose_stats *stats) {
    if (!c || !rgba || !pixels)
        return x3d::fail(X3D_ERR_INVALID_ARGUMENT, "null compositor, film or frame");
    if (int rc = check_size(width, height))
        return rc;
    const uint32_t npix = width * height;
    X3D_CHECK(c->film.ensure(static_cast<size_t>(npix) * 16));
    X3D_CHECK(c->frame.ensure(static_cast<size_t>(npix) * 4));
    X3D_CHECK(hipMemcpy(c->film.p, rgba, static_cast<size_t>(npix) * 16, hipMemcpyHostToDevice));
    if (int rc = commit_impl(c, static_cast<const float *>(c->film.p), npix, static_cast<uint8_t *>(c->frame.p), stats))
        return rc;
    X3D_CHECK(hipMemcpy(pixels, c->frame.p, static_cast<size_t>(npix) * 4, hipMemcpyDeviceToHost));
    return X3D_OK;
}

int x3d_compositor_commit_device(x3d_compositor *c, const float *d_rgba, uint32_t width, uint32_t height,
                                 uint8_t *d_pixels, x3d_compose_stats *stats) {
    if (!c || !d_rgba || !d_pixels)
        return x3d::fail(X3D_ERR_INVALID_ARGUMENT, "null compositor, film or frame");
    if (int rc = check_size(width, height))
        return rc;
    return commit_impl(c, d_rgba, width * height, d_pixels, stats);
}

} // extern "C"
