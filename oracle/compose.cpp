// ORACLE — test infrastructure only (see oracle.hpp). CPU restatement of the reference's film
// compositors (compositor.cpp:29-98), in two forms:
//
//   orc_compose       the library's arithmetic: its own double-precision log / exp (argument
//                     reduction + Taylor series, rounded once to float) and the fixed-order
//                     double sum of the exposure's logs (1024-pixel chunks: four sequential adds
//                     per lane of 256, pairwise tree; then the same over the chunk sums). The
//                     GPU frame must equal it byte for byte.
//   orc_compose_libm  the reference's arithmetic as written: std::pow / std::log / std::exp in
//                     float and the exposure's float sum in pixel order (compositor.cpp:81-98).
//                     The two forms differ only where a float result sits next to a rounding
//                     boundary (at most one step of the 8-bit output).
//
// The reference's compositor cannot be built here (compositor.h includes frame.h, which needs
// Qt), so both forms are pinned by closed forms in tests/test_oracle_compose.py: the gamma
// thresholds (k/255)^2.2, the ACES curve's values, the geometric-mean exposure, the counting
// rule of vec4 != 0 (tensor.h:15-17, 121-131); the log / exp / pow against float64 arithmetic.
// Conversions the reference leaves undefined (NaN, negative or >= 256 to unsigned char) are
// fixed here as in the library: NaN and negatives give 0, >= 255 gives 255.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr double kLn2Hi = 0x1.62e42feep-1;  // ln 2 to 32 significant bits
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;
constexpr float kInvGamma = 1.0f / 2.2f;  // if_compositor::pixel_of's exponent
constexpr uint32_t kLanes = 256, kChunk = 1024;

double from_bits(uint64_t b) {
    double x;
    std::memcpy(&x, &b, 8);
    return x;
}

uint64_t to_bits(double x) {
    uint64_t b;
    std::memcpy(&b, &x, 8);
    return b;
}

// log x = e ln 2 + 2 atanh((m - 1) / (m + 1)), x = m 2^e, m in [sqrt(1/2), sqrt(2)).
double log_d(double x) {
    if (!(x > 0.0))
        return x == 0.0 ? -INFINITY : NAN;
    if (x == INFINITY)
        return x;
    uint64_t b = to_bits(x);
    int e = static_cast<int>(b >> 52) - 1023;
    if (e == -1023) {
        b = to_bits(x * 0x1p54);
        e = static_cast<int>(b >> 52) - 1023 - 54;
    }
    double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > kSqrt2) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    for (int k = 10; k >= 0; k--)  // 1/21, 1/19, ..., 1/3, 1
        p = p * z + 1.0 / static_cast<double>(2 * k + 1);
    const double lm = (2.0 * s) * p;
    const double de = static_cast<double>(e);
    return de * kLn2Hi + (de * kLn2Lo + lm);
}

// e^t = 2^k e^r, t = k ln 2 + r, e^r by the Taylor polynomial of degree 13.
double exp_d(double t) {
    if (t != t)
        return t;
    if (t > 710.0)
        return INFINITY;
    if (t < -746.0)
        return 0.0;
    const double k = std::floor(t * kInvLn2 + 0.5);
    const double r = (t - k * kLn2Hi) - k * kLn2Lo;
    // 1/j! for j = 13 .. 0, each the correctly rounded quotient of exact integers
    static const double inv_fact[14] = {1.0,          1.0,           1.0 / 2.0,        1.0 / 6.0,
                                        1.0 / 24.0,   1.0 / 120.0,   1.0 / 720.0,      1.0 / 5040.0,
                                        1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0,
                                        1.0 / 479001600.0, 1.0 / 6227020800.0};
    double p = inv_fact[13];
    for (int j = 12; j >= 0; j--)
        p = p * r + inv_fact[j];
    const int ki = static_cast<int>(k);
    const int k1 = ki / 2, k2 = ki - k1;
    return p * from_bits(static_cast<uint64_t>(k1 + 1023) << 52) * from_bits(static_cast<uint64_t>(k2 + 1023) << 52);
}

float gamma_own(float x) {
    if (x != x)
        return x;
    if (x == 0.0f)
        return 0.0f;
    if (x == 1.0f)
        return 1.0f;
    return static_cast<float>(exp_d(static_cast<double>(kInvGamma) * log_d(static_cast<double>(x))));
}

float gamma_libm(float x) { return std::pow(x, kInvGamma); }

uint8_t to_byte(float v) {
    if (!(v > 0.0f))
        return 0;
    if (v >= 255.0f)
        return 255;
    return static_cast<uint8_t>(v);
}

// CLAMP (tensor.h:1003)
float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

// vec::at_least(0).at_most(1) (tensor.h:277-290)
float at_least0_at_most1(float x) {
    const float lo = 0.0f > x ? 0.0f : x;
    return 1.0f < lo ? 1.0f : lo;
}

// aces_compositor::aces_tonemap (compositor.cpp:68-79), one component
float aces(float x) {
    const float A = 2.51f, B = 0.03f, C = 2.43f, D = 0.59f, E = 0.14f;
    return x * (x * A + B) / (x * (x * C + D) + E);
}

// rgba_color != 0.0f (tensor.h:121-131 with equals(), tensor.h:15-17)
bool counted(float const *c) {
    for (int k = 0; k < 4; k++)
        if (!(std::abs(c[k] - 0.0f) <= 1e-5f))
            return true;
    return false;
}

// aces_compositor::luminance (compositor.cpp:77-79) + 1e-3
float lum_biased(float const *c) { return c[0] * 0.299f + c[1] * 0.587f + c[2] * 0.114f + 1e-3f; }

// Pairwise tree over kLanes values (v[t] += v[t + s], s = 128, 64, ..., 1).
double tree(std::vector<double> &v, std::vector<uint32_t> &n, uint32_t *count) {
    for (uint32_t s = kLanes / 2; s > 0; s >>= 1)
        for (uint32_t t = 0; t < s; t++) {
            v[t] = v[t] + v[t + s];
            n[t] = n[t] + n[t + s];
        }
    *count = n[0];
    return v[0];
}

float exposure_own(float const *rgba, uint32_t npix, uint32_t *count) {
    const uint32_t nchunks = (npix + kChunk - 1) / kChunk;
    std::vector<double> csum(nchunks), v(kLanes);
    std::vector<uint32_t> ccnt(nchunks), n(kLanes);
    for (uint32_t c = 0; c < nchunks; c++) {
        for (uint32_t t = 0; t < kLanes; t++) {
            double acc = 0.0;
            uint32_t k = 0;
            for (uint32_t j = 0; j < kChunk / kLanes; j++) {
                const uint32_t p = c * kChunk + j * kLanes + t;
                if (p < npix && counted(rgba + 4 * static_cast<size_t>(p))) {
                    acc = acc + static_cast<double>(static_cast<float>(log_d(lum_biased(rgba + 4 * static_cast<size_t>(p)))));
                    k++;
                }
            }
            v[t] = acc;
            n[t] = k;
        }
        csum[c] = tree(v, n, &ccnt[c]);
    }
    for (uint32_t t = 0; t < kLanes; t++) {
        double acc = 0.0;
        uint32_t k = 0;
        for (uint32_t c = t; c < nchunks; c += kLanes) {
            acc = acc + csum[c];
            k += ccnt[c];
        }
        v[t] = acc;
        n[t] = k;
    }
    const double sum = tree(v, n, count);
    const float mean = static_cast<float>(sum / static_cast<double>(*count));
    return static_cast<float>(exp_d(static_cast<double>(mean)));
}

// aces_compositor::exposure as written (compositor.cpp:85-98)
float exposure_libm(float const *rgba, uint32_t npix, uint32_t *count) {
    unsigned n = 0;
    float sum = 0;
    for (uint32_t i = 0; i < npix; i++)
        if (counted(rgba + 4 * static_cast<size_t>(i))) {
            sum += std::log(lum_biased(rgba + 4 * static_cast<size_t>(i)));
            n++;
        }
    sum /= n;
    *count = n;
    return std::exp(sum);
}

template <bool OWN>
void compose(float const *rgba, uint32_t w, uint32_t h, int kind, int auto_exposure, float exposure, uint8_t *out,
             float *exposure_used, uint32_t *counted_out) {
    const uint32_t npix = w * h;
    uint32_t cnt = 0;
    float e = exposure;
    if (kind == 1 && auto_exposure)
        e = OWN ? exposure_own(rgba, npix, &cnt) : exposure_libm(rgba, npix, &cnt);
    if (kind != 1)
        e = 1.0f;
    auto gamma = OWN ? gamma_own : gamma_libm;
    for (uint32_t p = 0; p < npix; p++) {
        float const *c = rgba + 4 * static_cast<size_t>(p);
        float r, g, b, a;
        if (kind == 1) {
            r = aces(c[0] * e);
            g = aces(c[1] * e);
            b = aces(c[2] * e);
            a = c[3];
        } else {
            r = at_least0_at_most1(c[0]);
            g = at_least0_at_most1(c[1]);
            b = at_least0_at_most1(c[2]);
            a = at_least0_at_most1(c[3]);
        }
        // if_compositor::pixel_of (compositor.cpp:29-39)
        uint8_t *o = out + 4 * static_cast<size_t>(p);
        o[0] = to_byte(gamma(clamp01(r)) * 255.0f);
        o[1] = to_byte(gamma(clamp01(g)) * 255.0f);
        o[2] = to_byte(gamma(clamp01(b)) * 255.0f);
        o[3] = to_byte(a * 255.0f);
    }
    if (exposure_used)
        *exposure_used = e;
    if (counted_out)
        *counted_out = cnt;
}

} // namespace

extern "C" {

// kind 0 = clamp_compositor, 1 = aces_compositor; out: uint8[4*w*h].
void orc_compose(float const *rgba, uint32_t w, uint32_t h, int kind, int auto_exposure, float exposure, uint8_t *out,
                 float *exposure_used, uint32_t *counted) {
    compose<true>(rgba, w, h, kind, auto_exposure, exposure, out, exposure_used, counted);
}

void orc_compose_libm(float const *rgba, uint32_t w, uint32_t h, int kind, int auto_exposure, float exposure,
                      uint8_t *out, float *exposure_used, uint32_t *counted) {
    compose<false>(rgba, w, h, kind, auto_exposure, exposure, out, exposure_used, counted);
}

// The compositor's own functions on n floats: fn 0 = pow(x, 1/2.2f), 1 = log, 2 = exp (each
// evaluated in double and rounded once to float).
void orc_compose_math(int fn, float const *x, uint32_t n, float *out) {
    for (uint32_t i = 0; i < n; i++) {
        if (fn == 0)
            out[i] = gamma_own(x[i]);
        else if (fn == 1)
            out[i] = static_cast<float>(log_d(static_cast<double>(x[i])));
        else
            out[i] = static_cast<float>(exp_d(static_cast<double>(x[i])));
    }
}

} // extern "C"
