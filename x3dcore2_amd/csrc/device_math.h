// Device-side math of the path-tracing hot path (gfx950, FP32).
//
// Operation order follows the reference renderer (DaviesX/x3dcore2, corelib/src) so that hit
// ids, hit distances and BSDF values are bit-identical to it; the library is compiled with
// -ffp-contract=off (no FMA contraction) and with IEEE division and square root.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scene_layout.h"

namespace x3d {

#define X3D_DI __device__ __forceinline__

// e8util::vec3 (tensor.h:32-308): scalar products are k*v(i); inner() accumulates from 0.
struct f3 {
    float x, y, z;
};
X3D_DI f3 mk(float x, float y, float z) { return f3{x, y, z}; }
X3D_DI f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
X3D_DI f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
X3D_DI f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
X3D_DI f3 operator*(float k, f3 a) { return f3{k * a.x, k * a.y, k * a.z}; }
X3D_DI f3 operator*(f3 a, float k) { return f3{k * a.x, k * a.y, k * a.z}; }
X3D_DI f3 operator/(f3 a, float k) { return f3{a.x / k, a.y / k, a.z / k}; }
X3D_DI f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
X3D_DI float dot(f3 a, f3 b) {
    float v = 0.0f;
    v += a.x * b.x;
    v += a.y * b.y;
    v += a.z * b.z;
    return v;
}
X3D_DI f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
X3D_DI float norm(f3 a) { return __builtin_sqrtf(dot(a, a)); }
X3D_DI f3 normalize(f3 a) { return (1.0f / norm(a)) * a; }
X3D_DI bool near0(float a) { return __builtin_fabsf(a - 0.0f) <= 1e-5f; }
X3D_DI bool near0(f3 a) { return near0(a.x) && near0(a.y) && near0(a.z); }
// libstdc++ std::max / std::min (NaN behaviour matters in the slab test).
X3D_DI float smax(float a, float b) { return (a < b) ? b : a; }
X3D_DI float smin(float a, float b) { return (b < a) ? b : a; }
X3D_DI f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

// A load through a pointer known to address device (global) memory. Scene arrays reached through
// pointers read from memory (a batch job's scene) are otherwise generic to the compiler, and a
// generic (flat) load counts against the LDS wait counter too, so every LDS stack access would
// wait for the node loads in flight.
typedef float x3d_v4f __attribute__((ext_vector_type(4)));
X3D_DI float4 ldg(const float4 *p) {
    // (through the builtin vector type: float4's copy constructor would load through a generic
    // reference again)
    const x3d_v4f v = *(const __attribute__((address_space(1))) x3d_v4f *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTMin = 1e-4f;   // pathspace.cpp:335
constexpr float kTMax = 1000.0f; // pathspace.cpp:336
// Relative slack of the distance cull (see traverse.h).
constexpr float kCullSlack = 1.0f / 4096.0f;

// Fast form of slab(): when none of the six plane distances is NaN, the per-axis updates reduce
// to one maximum of the entry distances and one minimum of the exit distances (min / max
// instructions, no compare-and-select pairs), and the reference's early exit after any axis
// equals one final comparison, because the entry distance only grows and the exit distance only
// shrinks from axis to axis. Values equal those of slab() (a zero's sign may differ, which no
// comparison sees). A distance (b - o) * vinv of finite floats is never NaN (overflow gives
// +-inf); with |b|, |o| <= 1e30 the difference is finite, so NaN needs a non-finite vinv: a zero
// (or denormal) direction component, i.e. a ray parallel to an axis, whose origin lies on a box
// plane (0 * inf). Rays for which that cannot be ruled out take slab() itself.
X3D_DI bool slab_may_nan(f3 o, f3 vinv, bool far_coords) {
    // (a zero vinv, from an infinite direction component, is excluded as well: it would turn
    // the infinite planes of an empty wide-node slot into NaN)
    return far_coords || !(__builtin_isfinite(vinv.x) && __builtin_isfinite(vinv.y) && __builtin_isfinite(vinv.z)) ||
           vinv.x == 0.0f || vinv.y == 0.0f || vinv.z == 0.0f ||
           !(fabsf(o.x) <= 1e30f && fabsf(o.y) <= 1e30f && fabsf(o.z) <= 1e30f);
}

X3D_DI bool slab_fast(float mnx, float mxx, float mny, float mxy, float mnz, float mxz, f3 o, f3 vinv,
                      float tmin, float tmax, float *t0) {
    float ax = (mnx - o.x) * vinv.x, bx = (mxx - o.x) * vinv.x;
    float ay = (mny - o.y) * vinv.y, by = (mxy - o.y) * vinv.y;
    float az = (mnz - o.z) * vinv.z, bz = (mxz - o.z) * vinv.z;
    float lo = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
    float hi = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
    *t0 = lo;
    return !(lo > hi);
}

// Slab test of aabb::intersect (tensor.h:815-840) on one box.
X3D_DI bool slab(float mnx, float mxx, float mny, float mxy, float mnz, float mxz, f3 o, f3 vinv,
                 float tmin, float tmax, float *t0) {
    // The reference returns as soon as tmin > tmax after an axis. Here all three axes are always
    // evaluated and `out` records whether such an exit would have happened, so the result is
    // the same (including NaN entry distances, which std::max / std::min propagate or drop in
    // argument order) while the loop body has no branches and all node loads issue up front.
    bool out = false;
    float k0 = (mnx - o.x) * vinv.x, k1 = (mxx - o.x) * vinv.x;
    bool sw = k0 > k1;
    tmin = smax(sw ? k1 : k0, tmin);
    tmax = smin(sw ? k0 : k1, tmax);
    out = out || tmin > tmax;
    k0 = (mny - o.y) * vinv.y;
    k1 = (mxy - o.y) * vinv.y;
    sw = k0 > k1;
    tmin = smax(sw ? k1 : k0, tmin);
    tmax = smin(sw ? k0 : k1, tmax);
    out = out || tmin > tmax;
    k0 = (mnz - o.z) * vinv.z;
    k1 = (mxz - o.z) * vinv.z;
    sw = k0 > k1;
    tmin = smax(sw ? k1 : k0, tmin);
    tmax = smin(sw ? k0 : k1, tmax);
    out = out || tmin > tmax;
    *t0 = tmin;
    return !out;
}

// ray::intersect (tensor.h:688-725), operation for operation.
X3D_DI bool tri_hit(f3 o, f3 v, f3 v0, f3 v1, f3 v2, float tmin, float tmax, float *b1, float *b2,
                    float *t) {
    f3 va = v1 - v0;
    f3 vb = v2 - v0;
    f3 vc = o - v0;
    f3 e2 = cross(v, vb);
    float det = dot(va, e2);
    if (det == 0.0f)
        return false;
    float inv = 1.0f / det;
    float c1 = dot(vc, e2) * inv;
    if (c1 < 0.0f || c1 > 1.0f)
        return false;
    f3 e1 = cross(vc, va);
    float c2 = dot(v, e1) * inv;
    if (c2 < 0.0f || c1 + c2 > 1.0f)
        return false;
    float tt = dot(vb, e1) * inv;
    if (tt < tmin || tt > tmax)
        return false;
    *b1 = c1;
    *b2 = c2;
    *t = tt;
    return true;
}

// ---------------------------------------------------------------------------------------------
// PCG32 (O'Neill 2014: pcg32_random_r / pcg32_srandom_r), one stream per (seed, pixel, sample).
// ---------------------------------------------------------------------------------------------
struct Pcg {
    uint64_t state, inc;
    X3D_DI uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    X3D_DI float draw() { return static_cast<float>(next() >> 8) * (1.0f / 16777216.0f); }
};
X3D_DI uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
X3D_DI Pcg pcg_for(uint64_t seed, uint64_t pixel, uint32_t sample) {
    Pcg g;
    g.state = 0;
    g.inc = (pixel << 1u) | 1u;
    g.next();
    g.state += splitmix64(seed * 0x9E3779B97F4A7C15ULL + sample);
    g.next();
    return g;
}

// sin(2 pi u), cos(2 pi u): quadrant reduction, Taylor polynomials of degree 9 / 10 on
// [-pi/4, pi/4], plain multiply-add chains (reproducible without FMA contraction).
X3D_DI void sincos_2pi(float u, float *s, float *c) {
    float v = 4.0f * u;
    float k = __builtin_rintf(v);
    float f = v - k;
    float r = f * 1.57079632679489661923f;
    float r2 = r * r;
    float sp = r + r * r2 * (-1.66666666666666666667e-1f +
                             r2 * (8.33333333333333333333e-3f +
                                   r2 * (-1.98412698412698412698e-4f +
                                         r2 * 2.75573192239858906526e-6f)));
    float cp = 1.0f + r2 * (-0.5f + r2 * (4.16666666666666666667e-2f +
                                          r2 * (-1.38888888888888888889e-3f +
                                                r2 * (2.48015873015873015873e-5f +
                                                      r2 * -2.75573192239858906526e-7f))));
    int q = static_cast<int>(k) & 3;
    float ss = q == 0 ? sp : (q == 1 ? cp : (q == 2 ? -sp : -cp));
    float cc = q == 0 ? cp : (q == 1 ? -sp : (q == 2 ? -cp : sp));
    *s = ss;
    *c = cc;
}

// vec3_basis (tensor.h:939-943).
X3D_DI void basis(f3 n, f3 *u, f3 *v) {
    f3 a = __builtin_fabsf(n.x) > .1f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 1.0f, 1.0f);
    *u = normalize(cross(a, n));
    *v = cross(n, *u);
}

// vec3_cos_hemisphere_sample (tensor.h:972-985).
X3D_DI f3 cos_hemisphere(f3 n, float e0, float e1) {
    f3 u, v;
    basis(n, &u, &v);
    float z = __builtin_sqrtf(e1);
    float r = __builtin_sqrtf(1.0f - z * z);
    float sp, cp;
    sincos_2pi(e0, &sp, &cp);
    float x = r * cp;
    float y = r * sp;
    return x * u + y * v + z * n;
}

// ---------------------------------------------------------------------------------------------
// Materials (material.cpp:10-39, 54-76, 78-106, 108-169, 171-252).
// ---------------------------------------------------------------------------------------------
X3D_DI float fresnel(DevMaterial const &m, f3 i, f3 h) {
    float re = m.ior_re, im = m.ior_im;
    float cf = 1.0f - dot(i, h);
    float cf2 = cf * cf;
    float cf4 = cf2 * cf2;
    float cf5 = cf4 * cf;
    float a = (re - 1) * (re - 1) + 4 * re * cf5 + im * im;
    float b = (re + 1) * (re + 1) + im * im;
    return a / b;
}
X3D_DI float ggx_d(float a2, f3 n, f3 h) {
    float ct = dot(n, h);
    float ct2 = ct * ct;
    float tan2 = (1.0f - ct2) / ct2;
    float c = 1.0f + tan2 / a2;
    return 1.0f / (kPi * a2 * ct2 * ct2 * c * c);
}
X3D_DI float ggx_lambda(float a2, f3 n, f3 w) {
    float ct = dot(n, w);
    float ct2 = ct * ct;
    float tan2 = (1.0f - ct2) / ct2;
    return (-1.0f + __builtin_sqrtf(1.0f + a2 * tan2)) / 2.0f;
}
X3D_DI float ggx_g(float a2, f3 i, f3 o, f3 n) {
    return 1.0f / (1.0f + ggx_lambda(a2, n, i) + ggx_lambda(a2, n, o));
}

// oren_nayar::eval factor k (material.cpp:134-161): the lobe is (albedo / pi) * k; false when
// either cosine is cut off.
X3D_DI bool oren_nayar_k(DevMaterial const &m, f3 n, f3 o, f3 i, float *k) {
    float ci = dot(i, n), co = dot(o, n);
    float cos_thei = ci < 0.0f ? 0.0f : (ci > 1.0f ? 1.0f : ci);
    float cos_theo = co < 0.0f ? 0.0f : (co > 1.0f ? 1.0f : co);
    if (near0(cos_thei) || near0(cos_theo))
        return false;
    float ca, cb;
    if (cos_thei < cos_theo) {
        ca = cos_thei;
        cb = cos_theo;
    } else {
        ca = cos_theo;
        cb = cos_thei;
    }
    float sin_a = __builtin_sqrtf(1.0f - ca * ca);
    float sin_b = __builtin_sqrtf(1.0f - cb * cb);
    float cos_io = ca * cb + sin_a * sin_b;
    float tan_b = sin_b / cos_theo;
    *k = m.on_a + m.on_b * smax(0.0f, cos_io) * sin_a * tan_b;
    return true;
}

// Evaluation of one non-mixture lobe.
X3D_DI f3 lobe_eval(DevMaterial const &m, f3 n, f3 o, f3 i) {
    f3 alb = mk(m.albedo[0], m.albedo[1], m.albedo[2]);
    if (m.kind == kMatFailSafe)
        return (1.0f / kPi) * alb;
    if (m.kind == kMatOrenNayar) {
        float k;
        if (!oren_nayar_k(m, n, o, i, &k))
            return mk(0.0f, 0.0f, 0.0f);
        return ((1.0f / kPi) * alb) * k;
    }
    // cook_torr
    float cos_o = smax(0.0f, dot(n, o));
    float cos_i = smax(0.0f, dot(n, i));
    if (near0(cos_i) || near0(cos_o))
        return mk(0.0f, 0.0f, 0.0f);
    f3 h = i + o;
    if (near0(h))
        return mk(0.0f, 0.0f, 0.0f);
    h = normalize(h);
    float f = fresnel(m, i, h);
    float d = ggx_d(m.alpha2, n, h);
    float g = ggx_g(m.alpha2, i, o, n);
    float c = f * d * g / (4.0f * cos_i * cos_o);
    return c * alb;
}

X3D_DI float lobe_pdf(DevMaterial const &m, f3 n, f3 o, f3 i) {
    if (m.kind != kMatCookTorr) {
        float c = dot(i, n);
        return c > 0.0f ? c / kPi : 0.0f;
    }
    if (dot(i, n) < 0.0f)
        return 0.0f;
    f3 h = i + o;
    if (near0(h))
        return 0.0f;
    h = normalize(h);
    float hdo = dot(h, o);
    if (!(hdo > 0.0f))
        return 0.0f;
    float ct = dot(n, h);
    if (!(ct > 0.0f))
        return 0.0f;
    return ggx_d(m.sample_alpha2, n, h) * ct / (4.0f * hdo);
}

X3D_DI f3 lobe_sample(DevMaterial const &m, f3 n, f3 o, Pcg &g, float *dens) {
    if (m.kind != kMatCookTorr) {
        float e0 = g.draw();
        float e1 = g.draw();
        f3 i = cos_hemisphere(n, e0, e1);
        *dens = dot(i, n) / kPi;
        return i;
    }
    float ms = m.sample_alpha2;
    f3 u, v;
    basis(n, &u, &v);
    float e0 = g.draw();
    float t = g.draw();
    float st, ct;
    sincos_2pi(e0, &st, &ct);
    float q = ms * t / (1.0f - t);
    float cos_phi = 1.0f / __builtin_sqrtf(1.0f + q);
    float sin_phi = __builtin_sqrtf(q) * cos_phi;
    float x = sin_phi * ct;
    float y = sin_phi * st;
    float z = cos_phi;
    f3 h = x * u + y * v + z * n;
    float hdo = dot(h, o);
    f3 i = (2.0f * hdo) * h - o;
    if (dot(i, n) < 0.0f) {
        *dens = 0.0f;
        return mk(0.0f, 0.0f, 0.0f);
    }
    *dens = ggx_d(ms, n, h) * cos_phi / (4.0f * hdo);
    return i;
}

// Material-level entry points; a mixture's operands are non-mixture materials.
X3D_DI f3 bsdf_eval(DevMaterial const *mats, int mat, f3 n, f3 o, f3 i) {
    DevMaterial const &m = mats[mat];
    if (m.kind != kMatMixture)
        return lobe_eval(m, n, o, i);
    float r = m.ratio;
    return r * lobe_eval(mats[m.sub0], n, o, i) + (1.0f - r) * lobe_eval(mats[m.sub1], n, o, i);
}
X3D_DI float bsdf_pdf(DevMaterial const *mats, int mat, f3 n, f3 o, f3 i) {
    DevMaterial const &m = mats[mat];
    if (m.kind != kMatMixture)
        return lobe_pdf(m, n, o, i);
    float r = m.ratio;
    return r * lobe_pdf(mats[m.sub0], n, o, i) + (1.0f - r) * lobe_pdf(mats[m.sub1], n, o, i);
}
// ref_mixture: the mixture density is the chosen lobe's density times its probability
// (mat_mixture::sample, material.cpp:94-106); otherwise the full mixture density.
X3D_DI f3 bsdf_sample(DevMaterial const *mats, int mat, f3 n, f3 o, Pcg &g, float *dens,
                      bool ref_mixture) {
    DevMaterial const &m = mats[mat];
    if (m.kind != kMatMixture)
        return lobe_sample(m, n, o, g, dens);
    float r = m.ratio;
    float sel = g.draw();
    f3 i;
    float d;
    if (sel < r) {
        i = lobe_sample(mats[m.sub0], n, o, g, &d);
        d = ref_mixture ? d * r : d;
    } else {
        i = lobe_sample(mats[m.sub1], n, o, g, &d);
        d = ref_mixture ? d * (1.0f - r) : d;
    }
    if (!ref_mixture && d != 0.0f)
        d = bsdf_pdf(mats, mat, n, o, i);
    *dens = d;
    return i;
}

} // namespace x3d
