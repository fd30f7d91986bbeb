// ORACLE — test infrastructure only.
//
// A CPU restatement of the reference renderer's hot path (DaviesX/x3dcore2, corelib/src), used
// as the checker for the HIP implementation in x3dcore2_amd/. Only tests/, __graft_entry__.smoke()
// and bench.py's CPU-baseline leg may load it. The product never links, imports or calls it.
//
// Every function cites the reference file:line it restates. The FP32 instantiation keeps the
// reference's operation order (no FMA contraction: built with -ffp-contract=off) so that BVH
// layout, closest-hit ids, hit t and BSDF eval can be compared bit-exactly. The FP64
// instantiation runs the same algorithms in double for high-accuracy image and gradient checks.
//
// Parity pins (tests/): BVH layout and hits against the reference's own code built into
// oracle/_ref (tests/test_oracle_ref.py) and against committed fixtures made from it
// (tests/golden/); brute force; closed forms (furnace, camera known answers of
// corelib/test/testcamera.cpp:33-49, BSDF reciprocity/energy of tst_material.cpp:54-106).

#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------------------------
// e8util::vec3 (corelib/src/tensor.h:32-308). inner() accumulates from 0 left to right
// (tensor.h:165-170); normalize() is (1/|v|)*v (tensor.h:225-227); scalar products are k*v(i).
// ---------------------------------------------------------------------------------------------
template <class T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T a) : x(a), y(a), z(a) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(T k, V3<T> a) { return {k * a.x, k * a.y, k * a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T k) { return {k * a.x, k * a.y, k * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T k) { return {a.x / k, a.y / k, a.z / k}; }
template <class T> inline V3<T> mod(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline T inner(V3<T> a, V3<T> b) {
    T v = 0;
    v += a.x * b.x;
    v += a.y * b.y;
    v += a.z * b.z;
    return v;
}
template <class T> inline V3<T> outer(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T norm(V3<T> a) { return std::sqrt(inner(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return (T(1.0f) / norm(a)) * a; }
// e8util::equals with T_CMP_EPSILON (tensor.h:14-16, 121-127).
template <class T> inline bool near0(T a) { return std::abs(a - T(0)) <= T(1e-5f); }
template <class T> inline bool near0(V3<T> a) { return near0(a.x) && near0(a.y) && near0(a.z); }
template <class U, class T> inline V3<U> cvt(V3<T> a) { return {U(a.x), U(a.y), U(a.z)}; }

// libstdc++ std::max / std::min, whose NaN behaviour the slab test depends on.
template <class T> inline T smax(T a, T b) { return (a < b) ? b : a; }
template <class T> inline T smin(T a, T b) { return (b < a) ? b : a; }

using V3f = V3<float>;

// ---------------------------------------------------------------------------------------------
// aabb (tensor.h:733-840), including its quirks: adding to a non-empty box uses the other
// operand's stored min/max even when that operand is empty (default min = max = 0).
// ---------------------------------------------------------------------------------------------
struct Box {
    V3f mn, mx;
    bool empty = true;
    Box() {}
    Box(V3f a, V3f b) : mn(a), mx(b) { empty = a.x > b.x || a.y > b.y || a.z > b.z; }
};
inline Box box_add(Box const &a, Box const &b) {
    if (!a.empty)
        return Box({smin(a.mn.x, b.mn.x), smin(a.mn.y, b.mn.y), smin(a.mn.z, b.mn.z)},
                   {smax(a.mx.x, b.mx.x), smax(a.mx.y, b.mx.y), smax(a.mx.z, b.mx.z)});
    return b;
}
inline Box box_add(Box const &a, V3f p) {
    if (!a.empty)
        return Box({smin(a.mn.x, p.x), smin(a.mn.y, p.y), smin(a.mn.z, p.z)},
                   {smax(a.mx.x, p.x), smax(a.mx.y, p.y), smax(a.mx.z, p.z)});
    return Box(p, p);
}
inline float surf_area(Box const &b) {
    if (b.empty)
        return 0.0f;
    V3f d = b.mx - b.mn;
    return 2.0f * (d.x * d.y + d.x * d.z + d.y * d.z);
}

// Slab test (tensor.h:815-840) in precision T on the stored FP32 bounds.
template <class T>
inline bool box_hit(Box const &b, V3<T> const &o, V3<T> const &vinv, T tmin, T tmax, T *t0) {
    for (int i = 0; i < 3; i++) {
        T k0 = (T(b.mn[i]) - o[i]) * vinv[i];
        T k1 = (T(b.mx[i]) - o[i]) * vinv[i];
        if (k0 > k1) {
            tmin = smax(k1, tmin);
            tmax = smin(k0, tmax);
        } else {
            tmin = smax(k0, tmin);
            tmax = smin(k1, tmax);
        }
        if (tmin > tmax)
            return false;
    }
    *t0 = tmin;
    return true;
}

// ray::intersect (tensor.h:688-725), op for op.
template <class T>
inline bool tri_hit(V3<T> const &o, V3<T> const &v, V3<T> const &v0, V3<T> const &v1,
                    V3<T> const &v2, T t_min, T t_max, T *b1, T *b2, T *t0) {
    V3<T> va = v1 - v0;
    V3<T> vb = v2 - v0;
    V3<T> vc = o - v0;
    V3<T> e2 = outer(v, vb);
    T det = inner(va, e2);
    if (det == T(0))
        return false;
    T inv = T(1.0f) / det;
    T bb1 = inner(vc, e2) * inv;
    if (bb1 < T(0) || bb1 > T(1))
        return false;
    V3<T> e1 = outer(vc, va);
    T bb2 = inner(v, e1) * inv;
    if (bb2 < T(0) || bb1 + bb2 > T(1))
        return false;
    T t = inner(vb, e1) * inv;
    if (t < t_min || t > t_max)
        return false;
    *b1 = bb1;
    *b2 = bb2;
    *t0 = t;
    return true;
}

// ---------------------------------------------------------------------------------------------
// Scene.
// ---------------------------------------------------------------------------------------------
enum MatKind : int { MAT_OREN_NAYAR = 0, MAT_COOK_TORR = 1, MAT_MIXTURE = 2 };

struct Material {
    int kind = MAT_OREN_NAYAR;
    V3f albedo{0.8f, 0.8f, 0.8f};
    float sigma = 0.0f;     // Oren-Nayar roughness (0 = Lambertian).
    float roughness = 0.0f; // cook_torr roughness r, alpha^2 = 2 r^2 (material.cpp:379).
    float sample_roughness = 0.0f; // roughness of the GGX sampling density (= roughness when added)
    float ior_re = 1.0f, ior_im = 0.0f;
    int sub0 = -1, sub1 = -1; // mixture operands
    float ratio = 0.5f;
};

struct Mesh {
    std::vector<V3f> verts, normals;
    std::vector<uint32_t> tris; // 3 per triangle
    int material = -1;          // -1: fail-safe (materialcontainer.cpp:7-17)
    int light = -1;
    std::vector<float> cum_area; // geometry.cpp:105-118
    float area = 0.0f;
};

// area_light (light.cpp:17-85) on mesh `mesh`, or sky_light (light.cpp:87-163) when mesh < 0.
// The sky's reference point and diameter are set at commit from the path space's bounds by
// sky_light::set_scene_boundary (light.cpp:96-100), which the reference never calls itself.
struct Light {
    int mesh;
    V3f rad;
    V3f power;
    bool sky() const { return mesh < 0; }
    V3f sky_ref;        // m_ref_p: the bounds' centroid with z = 0
    float sky_dia = 0;  // m_dia: twice the bounds' enclosing radius
};

struct FlatNode { // bvh_path_space_layout::flattened_node (pathspace.h:130-150)
    Box bound;
    uint32_t num_prims = 0, split_axis = 0xFF, next_child = 0xFFFFFFFFu, prim_start = 0xFFFFFFFFu;
};

struct Prim {
    uint32_t geo, tri;
};

struct Scene {
    std::vector<Material> mats;
    std::vector<Mesh> meshes;
    std::vector<Light> lights;
    std::vector<float> light_cum; // basic_light_sources::commit (lightsources.cpp:45-52)
    float total_power = 0.0f;
    std::vector<FlatNode> nodes;
    std::vector<Prim> prims;
    // per-prim vertex positions cached for traversal
    std::vector<V3f> pv0, pv1, pv2;
    // if_path_space::aabb (pathspace.cpp:13-18): lower corner and max - min of all vertices
    V3f space_lo, space_range;
};

// BVH build (pathspace.cpp:121-328): SAH with 20 buckets, <=4 prims per leaf, median split by
// std::sort when depth > log2(N), std::partition otherwise, pre-order flatten.
void build_bvh(Scene &s);
void commit(Scene &s);

// ---------------------------------------------------------------------------------------------
// Hits.
// ---------------------------------------------------------------------------------------------
template <class T> struct Hit {
    int32_t prim = -1; // index into Scene::prims; -1 = miss
    T t = std::numeric_limits<T>::infinity();
    T b1 = 0, b2 = 0;
    V3<T> p, n; // vertex and shading normal (pathspace.cpp:385-409)
    int geo = -1;
    bool valid() const { return prim >= 0; }
};

// Traversal order. REF: the reference's LIFO order without distance culling
// (pathspace.cpp:330-413). CULLED: near-child-first with a conservative distance cull and the
// reference's tie rule made order independent (equal t: larger leaf prim_start wins, then lower
// prim index), which is what the GPU kernels implement.
enum Traversal : int { TRAV_REF = 0, TRAV_CULLED = 1, TRAV_BRUTE = 2 };

// Relative slack of the distance cull: a box is skipped only when its entry distance exceeds
// the best hit by more than this fraction, so a hit whose FP32 t is computed slightly below a
// box's entry distance is never lost.
constexpr float kCullSlack = 1.0f / 4096.0f;

template <class T> Hit<T> intersect(Scene const &s, V3<T> o, V3<T> v, int mode, uint64_t *stats);
template <class T>
bool occluded(Scene const &s, V3<T> o, V3<T> v, T tmin, T tmax, int mode, uint64_t *stats);

// ---------------------------------------------------------------------------------------------
// Random numbers: PCG32 (O'Neill 2014, pcg32_random_r / pcg32_srandom_r), one stream per
// (seed, pixel, sample). Replaces the reference's shared minstd_rand0 stream (tensor.h:987-1001),
// which cannot be split across GPU threads.
// ---------------------------------------------------------------------------------------------
struct Pcg {
    uint64_t state, inc;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    float draw() { return static_cast<float>(next() >> 8) * (1.0f / 16777216.0f); }
};
inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline Pcg pcg_for(uint64_t seed, uint64_t pixel, uint32_t sample) {
    Pcg g;
    g.state = 0;
    g.inc = (pixel << 1u) | 1u;
    g.next();
    g.state += splitmix64(seed * 0x9E3779B97F4A7C15ULL + sample);
    g.next();
    return g;
}

// sin(2 pi u), cos(2 pi u) for u in [0, 1): quadrant reduction plus Taylor polynomials on
// [-pi/4, pi/4] (degree 9 / 10), evaluated with plain multiplies and adds so that the FP32
// result is reproducible on any IEEE machine without FMA contraction.
void sincos_2pi(float u, float *s, float *c);
void sincos_2pi(double u, double *s, double *c);

// ---------------------------------------------------------------------------------------------
// Render options and entry points.
// ---------------------------------------------------------------------------------------------
enum Tracer : int {
    TR_NORMAL = 0, TR_POSITION = 1, TR_DIRECT = 2, TR_UNIDIRECT = 3, TR_UNIDIRECT_LT1 = 4,
    TR_BIDIRECT_LT2 = 5, TR_BIDIRECT_MIS = 6, TR_UNIDIRECT_MIS = 16
};

struct Camera {
    float sensor, focal, aspect;
    float m[16]; // column-major 4x4 transform; camera.cpp:63-71 takes R = m(0..2,0..2), t = m(0..2,3)
};

struct RenderOpts {
    int tracer = TR_UNIDIRECT_MIS;
    uint32_t max_depth = 5;     // bounces after the first hit (config mode)
    uint32_t max_path_len = 8;  // BDPT subpath vertices (pathtracer.h:201)
    uint64_t seed = 1;
    int traversal = TRAV_CULLED;
};

void camera_ray(Camera const &c, uint32_t x, uint32_t y, uint32_t w, uint32_t h, V3f *o, V3f *v);

// Radiance estimate of one sample of one pixel.
template <class T>
V3<T> trace_sample(Scene const &s, RenderOpts const &opt, Hit<T> const &first, V3<T> const &ray_v,
                   uint64_t pixel, uint32_t sample, uint64_t *stats);

} // namespace orc
