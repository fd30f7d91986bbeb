// ORACLE — test infrastructure only (see oracle.hpp). C++ restatement of the reference's hot path.

#include "oracle.hpp"

#include <array>
#include <atomic>
#include <thread>

namespace orc {

// ---------------------------------------------------------------------------------------------
// Scene commit: per-mesh area CDF (geometry.cpp:105-118), light power CDF
// (light.cpp:20, lightsources.cpp:45-52), BVH (pathspace.cpp:292-328).
// ---------------------------------------------------------------------------------------------
void commit(Scene &s) {
    for (Mesh &m : s.meshes) {
        size_t nt = m.tris.size() / 3;
        m.cum_area.resize(nt);
        float cum = 0;
        for (size_t i = 0; i < nt; i++) {
            V3f v0 = m.verts[m.tris[3 * i]], v1 = m.verts[m.tris[3 * i + 1]],
                v2 = m.verts[m.tris[3 * i + 2]];
            float a = 0.5f * norm(outer(v1 - v0, v2 - v0));
            cum += a;
            m.cum_area[i] = cum;
        }
        m.area = cum;
        m.light = -1;
    }
    // if_path_space::load unions each mesh's aabb (trimesh::update_aabb, geometry.cpp:97-103)
    // into an initially empty box (tensor.h:756-771)
    bool empty = true;
    V3f lo, hi;
    for (Mesh const &m : s.meshes)
        for (V3f const &v : m.verts) {
            lo = empty ? v : V3f(std::min(lo.x, v.x), std::min(lo.y, v.y), std::min(lo.z, v.z));
            hi = empty ? v : V3f(std::max(hi.x, v.x), std::max(hi.y, v.y), std::max(hi.z, v.z));
            empty = false;
        }
    s.space_lo = lo;
    s.space_range = hi - lo;
    s.light_cum.clear();
    s.total_power = 0;
    for (size_t l = 0; l < s.lights.size(); l++) {
        Light &L = s.lights[l];
        if (L.sky()) {
            // sky_light::set_scene_boundary (light.cpp:96-100) with the path space's bounds:
            // enclosing_radius = |max - min| / 2, centroid = (max + min) / 2 (tensor.h:811-813)
            L.sky_dia = 2 * (norm(hi - lo) / 2);
            L.sky_ref = (hi + lo) / 2.0f;
            L.sky_ref.z = 0.0f;
            // sky_light::power (light.cpp:151): pi * dia^2 / 2 in every channel
            L.power = V3f(static_cast<float>(M_PI) * (L.sky_dia * L.sky_dia / 2));
        } else {
            L.power = (static_cast<float>(M_PI) * s.meshes[L.mesh].area) * L.rad;
            s.meshes[L.mesh].light = static_cast<int>(l);
        }
        s.total_power += norm(L.power);
        s.light_cum.push_back(s.total_power);
    }
    build_bvh(s);
}

namespace {

struct PrimDetail {
    Prim prim;
    Box bound;
    V3f centroid;
};

struct Bucket {
    float cost = 0.0f;
    unsigned num_prims = 0;
    Box bound;
};

struct Node {
    Box bound;
    Node *child[2] = {nullptr, nullptr};
    uint32_t split_axis = 0xFF, prim_start = 0, num_prims = 0;
};

constexpr unsigned kMaxPrims = 4, kBuckets = 20, kTriCost = 8, kBoxCost = 1;

Box bound_of(std::vector<PrimDetail> const &p, unsigned a, unsigned b) { // pathspace.cpp:168-175
    Box r;
    for (unsigned i = a; i < b; i++)
        r = box_add(r, p[i].bound);
    return r;
}

// pathspace.cpp:130-166
std::vector<Bucket> sah_buckets(std::vector<PrimDetail> const &p, unsigned start, unsigned end,
                                unsigned axis, Box const &b, V3f range) {
    std::vector<Bucket> bk(kBuckets);
    float width = range[axis] / kBuckets;
    for (unsigned i = start; i < end; i++) {
        float q = (p[i].centroid[axis] - b.mn[axis]) / width;
        // q is NaN only when every centroid on the widest axis coincides (width == 0); the
        // reference's float-to-unsigned conversion then yields bucket 0 on x86-64.
        unsigned ib = (q != q) ? 0u : static_cast<unsigned>(q);
        if (ib == kBuckets)
            ib = kBuckets - 1;
        bk[ib].bound = box_add(bk[ib].bound, p[i].bound);
        bk[ib].num_prims++;
    }
    bk[kBuckets - 1].cost = INFINITY;
    for (unsigned i = 0; i < kBuckets - 1; i++) {
        Box first, second;
        unsigned cf = 0, cs = 0;
        for (unsigned j = 0; j <= i; j++) {
            first = box_add(first, bk[j].bound);
            cf += bk[j].num_prims;
        }
        for (unsigned j = i + 1; j < kBuckets; j++) {
            second = box_add(second, bk[j].bound);
            cs += bk[j].num_prims;
        }
        bk[i].cost = kBoxCost + kTriCost * (+surf_area(first) / surf_area(b) * cf +
                                            surf_area(second) / surf_area(b) * cs);
    }
    return bk;
}

// pathspace.cpp:177-260
Node *build(std::vector<PrimDetail> &p, unsigned start, unsigned end, unsigned depth) {
    if (end - start == 0)
        return nullptr;
    if (end - start == 1) {
        Node *n = new Node();
        n->bound = p[start].bound;
        n->prim_start = start;
        n->num_prims = 1;
        return n;
    }
    Box b = bound_of(p, start, end);
    V3f range = b.mx - b.mn;
    unsigned axis;
    if (range.x > range.y && range.x > range.z)
        axis = 0;
    else if (range.y > range.z)
        axis = 1;
    else
        axis = 2;
    float cost_nonsplit = (end - start) * kTriCost;
    std::vector<Bucket> bk = sah_buckets(p, start, end, axis, b, range);
    unsigned split_antiaxis = 1;
    float cost_split = bk[0].cost;
    for (unsigned i = 1; i < bk.size() - 1; i++) {
        if (bk[i].cost < cost_split) {
            split_antiaxis = i;
            cost_split = bk[i].cost;
        }
    }
    if (cost_split < cost_nonsplit || end - start > kMaxPrims) {
        unsigned mid;
        if (depth > std::log2(p.size())) {
            std::sort(p.begin() + start, p.begin() + end,
                      [axis](PrimDetail const &a, PrimDetail const &c) {
                          return a.centroid[axis] < c.centroid[axis];
                      });
            mid = (start + end) >> 1;
        } else {
            float split = b.mn[axis] + (split_antiaxis + 1) * range[axis] / kBuckets;
            auto it = std::partition(p.begin() + start, p.begin() + end,
                                     [axis, split](PrimDetail const &a) {
                                         return a.centroid[axis] < split;
                                     });
            mid = static_cast<unsigned>(it - p.begin());
            if (mid == start)
                mid++;
            else if (mid == end)
                mid--;
        }
        Node *n = new Node();
        n->bound = b;
        n->split_axis = axis;
        n->child[0] = build(p, start, mid, depth + 1);
        n->child[1] = build(p, mid, end, depth + 1);
        return n;
    }
    Node *n = new Node();
    n->bound = b;
    n->prim_start = start;
    n->num_prims = end - start;
    return n;
}

// pathspace.cpp:274-290
void flatten(std::vector<FlatNode> &out, Node *n) {
    if (!n)
        return;
    if (n->num_prims > 0) {
        FlatNode f;
        f.bound = n->bound;
        f.num_prims = n->num_prims;
        f.split_axis = 0xFF;
        f.next_child = 0xFFFFFFFFu;
        f.prim_start = n->prim_start;
        out.push_back(f);
    } else {
        size_t p = out.size();
        out.push_back(FlatNode());
        flatten(out, n->child[0]);
        FlatNode f;
        f.bound = n->bound;
        f.num_prims = 0;
        f.split_axis = n->split_axis;
        f.next_child = static_cast<uint32_t>(out.size());
        f.prim_start = 0xFFFFFFFFu;
        out[p] = f;
        flatten(out, n->child[1]);
    }
}

void free_tree(Node *n) {
    if (!n)
        return;
    free_tree(n->child[0]);
    free_tree(n->child[1]);
    delete n;
}

} // namespace

void build_bvh(Scene &s) {
    std::vector<PrimDetail> p;
    for (uint32_t g = 0; g < s.meshes.size(); g++) {
        Mesh const &m = s.meshes[g];
        for (uint32_t t = 0; t < m.tris.size() / 3; t++) {
            PrimDetail d; // pathspace.h:97-113
            d.prim = {g, t};
            d.bound = box_add(d.bound, m.verts[m.tris[3 * t]]);
            d.bound = box_add(d.bound, m.verts[m.tris[3 * t + 1]]);
            d.bound = box_add(d.bound, m.verts[m.tris[3 * t + 2]]);
            d.centroid = (d.bound.mx + d.bound.mn) / 2.0f;
            p.push_back(d);
        }
    }
    Node *root = build(p, 0, static_cast<unsigned>(p.size()), 0);
    s.nodes.clear();
    flatten(s.nodes, root);
    free_tree(root);
    s.prims.resize(p.size());
    s.pv0.resize(p.size());
    s.pv1.resize(p.size());
    s.pv2.resize(p.size());
    for (size_t i = 0; i < p.size(); i++) {
        s.prims[i] = p[i].prim;
        Mesh const &m = s.meshes[p[i].prim.geo];
        s.pv0[i] = m.verts[m.tris[3 * p[i].prim.tri]];
        s.pv1[i] = m.verts[m.tris[3 * p[i].prim.tri + 1]];
        s.pv2[i] = m.verts[m.tris[3 * p[i].prim.tri + 2]];
    }
}

// ---------------------------------------------------------------------------------------------
// Traversal.
// ---------------------------------------------------------------------------------------------
namespace {

template <class T> void finish_hit(Scene const &s, Hit<T> &h) { // pathspace.cpp:385-409
    Prim pr = s.prims[h.prim];
    Mesh const &m = s.meshes[pr.geo];
    uint32_t i0 = m.tris[3 * pr.tri], i1 = m.tris[3 * pr.tri + 1], i2 = m.tris[3 * pr.tri + 2];
    T b0 = T(1.0f) - h.b1 - h.b2;
    h.p = b0 * cvt<T>(m.verts[i0]) + h.b1 * cvt<T>(m.verts[i1]) + h.b2 * cvt<T>(m.verts[i2]);
    h.n = normalize(b0 * cvt<T>(m.normals[i0]) + h.b1 * cvt<T>(m.normals[i1]) +
                    h.b2 * cvt<T>(m.normals[i2]));
    h.geo = static_cast<int>(pr.geo);
}

template <class T>
inline bool test_prim(Scene const &s, uint32_t i, V3<T> const &o, V3<T> const &v, T tmin, T tmax,
                      T *b1, T *b2, T *t) {
    return tri_hit(o, v, cvt<T>(s.pv0[i]), cvt<T>(s.pv1[i]), cvt<T>(s.pv2[i]), tmin, tmax, b1, b2,
                   t);
}

} // namespace

template <class T> Hit<T> intersect(Scene const &s, V3<T> o, V3<T> v, int mode, uint64_t *stats) {
    Hit<T> h;
    if (s.nodes.empty())
        return h;
    T const tmin = T(1e-4f), tmax = T(1000.0f);
    V3<T> vinv = {T(1.0f) / v.x, T(1.0f) / v.y, T(1.0f) / v.z};
    T t = std::numeric_limits<T>::infinity();
    T b1 = 0, b2 = 0;
    int32_t best = -1;
    uint32_t best_leaf = 0;
    uint64_t pops = 0, tests = 0;

    if (mode == TRAV_BRUTE) { // linear_path_space_layout::intersect order (pathspace.cpp:36-92)
        // Geometry-major, triangle-minor order; the first triangle reaching the minimum wins.
        std::vector<uint32_t> order(s.prims.size());
        for (uint32_t i = 0; i < s.prims.size(); i++)
            order[i] = i;
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
            return s.prims[a].geo != s.prims[c].geo ? s.prims[a].geo < s.prims[c].geo
                                                    : s.prims[a].tri < s.prims[c].tri;
        });
        for (uint32_t i : order) {
            T c1, c2, t0;
            tests++;
            if (test_prim(s, i, o, v, tmin, tmax, &c1, &c2, &t0) && t0 < t) {
                t = t0;
                b1 = c1;
                b2 = c2;
                best = static_cast<int32_t>(i);
            }
        }
    } else if (mode == TRAV_REF) { // pathspace.cpp:343-383
        std::vector<uint32_t> stack{0};
        while (!stack.empty()) {
            uint32_t n = stack.back();
            stack.pop_back();
            FlatNode const &nd = s.nodes[n];
            if (nd.num_prims > 0) {
                for (uint32_t i = nd.prim_start; i < nd.prim_start + nd.num_prims; i++) {
                    T c1, c2, t0;
                    tests++;
                    if (test_prim(s, i, o, v, tmin, tmax, &c1, &c2, &t0) && t0 < t) {
                        t = t0;
                        b1 = c1;
                        b2 = c2;
                        best = static_cast<int32_t>(i);
                    }
                }
            } else {
                pops++;
                T t0;
                if (box_hit(s.nodes[n + 1].bound, o, vinv, tmin, tmax, &t0))
                    stack.push_back(n + 1);
                if (box_hit(s.nodes[nd.next_child].bound, o, vinv, tmin, tmax, &t0))
                    stack.push_back(nd.next_child);
            }
        }
    } else { // TRAV_CULLED
        struct E {
            uint32_t n;
            T tent;
        };
        std::vector<E> stack{{0, T(0)}};
        T const slack = T(kCullSlack);
        while (!stack.empty()) {
            E e = stack.back();
            stack.pop_back();
            if (e.tent > t + t * slack)
                continue;
            FlatNode const &nd = s.nodes[e.n];
            if (nd.num_prims > 0) {
                for (uint32_t i = nd.prim_start; i < nd.prim_start + nd.num_prims; i++) {
                    T c1, c2, t0;
                    tests++;
                    if (test_prim(s, i, o, v, tmin, tmax, &c1, &c2, &t0) &&
                        (t0 < t || (t0 == t && nd.prim_start > best_leaf))) {
                        t = t0;
                        b1 = c1;
                        b2 = c2;
                        best = static_cast<int32_t>(i);
                        best_leaf = nd.prim_start;
                    }
                }
            } else {
                pops++;
                T tl = 0, tr = 0;
                T cut = t + t * slack;
                bool hl = box_hit(s.nodes[e.n + 1].bound, o, vinv, tmin, tmax, &tl) && !(tl > cut);
                bool hr = box_hit(s.nodes[nd.next_child].bound, o, vinv, tmin, tmax, &tr) &&
                          !(tr > cut);
                if (hl && hr) {
                    if (tr < tl) {
                        stack.push_back({e.n + 1, tl});
                        stack.push_back({nd.next_child, tr});
                    } else {
                        stack.push_back({nd.next_child, tr});
                        stack.push_back({e.n + 1, tl});
                    }
                } else if (hl) {
                    stack.push_back({e.n + 1, tl});
                } else if (hr) {
                    stack.push_back({nd.next_child, tr});
                }
            }
        }
    }
    if (stats) {
        stats[0] += pops;
        stats[1] += tests;
    }
    if (best >= 0) {
        h.prim = best;
        h.t = t;
        h.b1 = b1;
        h.b2 = b2;
        finish_hit(s, h);
    }
    return h;
}

// has_intersect (pathspace.cpp:415-456): any triangle in (tmin, tmax) under boxes clipped to the
// same interval. The answer does not depend on the visiting order.
template <class T>
bool occluded(Scene const &s, V3<T> o, V3<T> v, T tmin, T tmax, int mode, uint64_t *stats) {
    if (s.nodes.empty())
        return false;
    V3<T> vinv = {T(1.0f) / v.x, T(1.0f) / v.y, T(1.0f) / v.z};
    uint64_t pops = 0, tests = 0;
    bool hit = false;
    if (mode == TRAV_BRUTE) {
        for (uint32_t i = 0; i < s.prims.size() && !hit; i++) {
            T c1, c2, t0;
            tests++;
            hit = test_prim(s, i, o, v, tmin, tmax, &c1, &c2, &t0);
        }
    } else {
        std::vector<uint32_t> stack{0};
        while (!stack.empty() && !hit) {
            uint32_t n = stack.back();
            stack.pop_back();
            FlatNode const &nd = s.nodes[n];
            if (nd.num_prims > 0) {
                for (uint32_t i = nd.prim_start; i < nd.prim_start + nd.num_prims; i++) {
                    T c1, c2, t0;
                    tests++;
                    if (test_prim(s, i, o, v, tmin, tmax, &c1, &c2, &t0)) {
                        hit = true;
                        break;
                    }
                }
            } else {
                pops++;
                T t0;
                if (box_hit(s.nodes[n + 1].bound, o, vinv, tmin, tmax, &t0))
                    stack.push_back(n + 1);
                if (box_hit(s.nodes[nd.next_child].bound, o, vinv, tmin, tmax, &t0))
                    stack.push_back(nd.next_child);
            }
        }
    }
    if (stats) {
        stats[2] += pops;
        stats[3] += tests;
    }
    return hit;
}

template Hit<float> intersect<float>(Scene const &, V3f, V3f, int, uint64_t *);
template Hit<double> intersect<double>(Scene const &, V3<double>, V3<double>, int, uint64_t *);
template bool occluded<float>(Scene const &, V3f, V3f, float, float, int, uint64_t *);
template bool occluded<double>(Scene const &, V3<double>, V3<double>, double, double, int,
                               uint64_t *);

// ---------------------------------------------------------------------------------------------
// sin/cos of 2 pi u.
// ---------------------------------------------------------------------------------------------
void sincos_2pi(float u, float *s, float *c) {
    float v = 4.0f * u;
    float k = std::nearbyint(v);
    float f = v - k; // exact, |f| <= 0.5
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
    switch (q) {
    case 0: *s = sp; *c = cp; break;
    case 1: *s = cp; *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
    }
}
void sincos_2pi(double u, double *s, double *c) {
    double a = 2.0 * M_PI * u;
    *s = std::sin(a);
    *c = std::cos(a);
}

// ---------------------------------------------------------------------------------------------
// Camera (camera.cpp:38-48 after camera.cpp:63-71).
// ---------------------------------------------------------------------------------------------
void camera_ray(Camera const &c, uint32_t x, uint32_t y, uint32_t w, uint32_t h, V3f *o, V3f *v) {
    float s = c.sensor;
    V3f d((static_cast<float>(x) / (w - 1) * s - s / 2.0f) * c.aspect,
          (h - 1 - static_cast<float>(y)) / (h - 1) * s - s / 2.0f, -c.focal);
    // m_r * v.homo(0): accumulate over columns j then rows i from zero (tensor.h:421-430).
    V3f r;
    for (int i = 0; i < 3; i++) {
        float a = 0;
        a = a + c.m[0 * 4 + i] * d.x;
        a = a + c.m[1 * 4 + i] * d.y;
        a = a + c.m[2 * 4 + i] * d.z;
        a = a + 0.0f * 0.0f;
        r[i] = a;
    }
    *o = V3f(c.m[12], c.m[13], c.m[14]);
    *v = normalize(r);
}

// ---------------------------------------------------------------------------------------------
// Sampling helpers (tensor.h:939-985).
// ---------------------------------------------------------------------------------------------
namespace {

template <class T> void basis(V3<T> n, V3<T> *u, V3<T> *v) { // tensor.h:939-943
    V3<T> a = std::abs(n.x) > T(.1f) ? V3<T>(0, 1, 0) : V3<T>(1, 1, 1);
    *u = normalize(outer(a, n));
    *v = outer(n, *u);
}

template <class T> V3<T> cos_hemisphere(V3<T> n, T e0, T e1) { // tensor.h:972-985
    V3<T> u, v;
    basis(n, &u, &v);
    T z = std::sqrt(e1);
    T r = std::sqrt(T(1) - z * z);
    T sp, cp;
    sincos_2pi(e0, &sp, &cp);
    T x = r * cp;
    T y = r * sp;
    return x * u + y * v + z * n;
}

template <class T> T pi_v() { return T(M_PI); }

// ---------------------------------------------------------------------------------------------
// Materials (material.cpp:10-39, 54-76, 78-106, 108-169, 171-252).
// ---------------------------------------------------------------------------------------------
template <class T> T fresnel(Material const &m, V3<T> i, V3<T> h) {
    T re = T(m.ior_re), im = T(m.ior_im);
    T cf = T(1.0f) - inner(i, h);
    T cf2 = cf * cf;
    T cf4 = cf2 * cf2;
    T cf5 = cf4 * cf;
    T a = (re - 1) * (re - 1) + 4 * re * cf5 + im * im;
    T b = (re + 1) * (re + 1) + im * im;
    return a / b;
}
template <class T> T ggx_d(T a2, V3<T> n, V3<T> h) {
    T ct = inner(n, h);
    T ct2 = ct * ct;
    T tan2 = (T(1.0f) - ct2) / ct2;
    T c = T(1.0f) + tan2 / a2;
    return T(1.0f) / (pi_v<T>() * a2 * ct2 * ct2 * c * c);
}
template <class T> T ggx_lambda(T a2, V3<T> n, V3<T> w) {
    T ct = inner(n, w);
    T ct2 = ct * ct;
    T tan2 = (T(1.0f) - ct2) / ct2;
    return (T(-1.0f) + std::sqrt(T(1.0f) + a2 * tan2)) / T(2.0f);
}
template <class T> T ggx_g(T a2, V3<T> i, V3<T> o, V3<T> n) {
    return T(1.0f) / (T(1.0f) + ggx_lambda(a2, n, i) + ggx_lambda(a2, n, o));
}
template <class T> T alpha2_of(Material const &m) {
    // float in the FP32 build: m_alpha2(2 * roughness * roughness) (material.cpp:379)
    return T(2) * T(m.roughness) * T(m.roughness);
}
// alpha^2 of the GGX sampling density (and its pdf): the evaluated roughness unless the
// evaluated parameters were changed with sampling left alone (orc_set_material_params).
template <class T> T sample_alpha2_of(Material const &m) {
    return T(2) * T(m.sample_roughness) * T(m.sample_roughness);
}

template <class T> V3<T> clampv(T x) { return V3<T>(x); }

} // namespace

// oren_nayar::eval factor k (material.cpp:134-161): the lobe is (albedo / pi) * k; false when
// either cosine is cut off.
template <class T> bool oren_nayar_k(Material const &m, V3<T> n, V3<T> o, V3<T> i, T *k_out) {
    T sg = T(m.sigma);
    T sigma2 = sg * sg;
    T A = T(1) - T(0.5f) * sigma2 / (sigma2 + T(0.33f));
    T B = T(0.45f) * sigma2 / (sigma2 + T(0.09f));
    T ci = inner(i, n), co = inner(o, n);
    T cos_thei = ci < T(0) ? T(0) : (ci > T(1) ? T(1) : ci);
    T cos_theo = co < T(0) ? T(0) : (co > T(1) ? T(1) : co);
    if (near0(cos_thei) || near0(cos_theo))
        return false;
    T ca, cb;
    if (cos_thei < cos_theo) {
        ca = cos_thei;
        cb = cos_theo;
    } else {
        ca = cos_theo;
        cb = cos_thei;
    }
    T sin_a = std::sqrt(T(1.0f) - ca * ca);
    T sin_b = std::sqrt(T(1.0f) - cb * cb);
    T cos_io = ca * cb + sin_a * sin_b;
    T tan_b = sin_b / cos_theo;
    *k_out = A + B * smax(T(0.0f), cos_io) * sin_a * tan_b;
    return true;
}

template <class T> V3<T> bsdf_eval(Scene const &s, int mat, V3<T> n, V3<T> o, V3<T> i) {
    if (mat < 0) // mat_fail_safe::eval (material.cpp:64-68)
        return (T(1.0f) / pi_v<T>()) * V3<T>(T(0.8f), T(0.8f), T(0.8f));
    Material const &m = s.mats[mat];
    V3<T> alb = cvt<T>(m.albedo);
    if (m.kind == MAT_OREN_NAYAR) { // material.cpp:134-161
        T k;
        if (!oren_nayar_k<T>(m, n, o, i, &k))
            return V3<T>();
        return ((T(1.0f) / pi_v<T>()) * alb) * k;
    }
    if (m.kind == MAT_COOK_TORR) { // material.cpp:171-198
        T cos_o = smax(T(0.0f), inner(n, o));
        T cos_i = smax(T(0.0f), inner(n, i));
        if (near0(cos_i) || near0(cos_o))
            return V3<T>();
        V3<T> h = i + o;
        if (near0(h))
            return V3<T>();
        h = normalize(h);
        T a2 = alpha2_of<T>(m);
        T f = fresnel(m, i, h);
        T d = ggx_d(a2, n, h);
        T g = ggx_g(a2, i, o, n);
        T c = f * d * g / (T(4.0f) * cos_i * cos_o);
        return c * alb;
    }
    // mat_mixture::eval (material.cpp:89-92)
    T r = T(m.ratio);
    return r * bsdf_eval<T>(s, m.sub0, n, o, i) + (T(1) - r) * bsdf_eval<T>(s, m.sub1, n, o, i);
}

// Solid-angle density with which bsdf_sample would return direction i (config-mode MIS; the
// reference has no pdf(), its sample() densities are restated here as functions of (o, i)).
template <class T> T bsdf_pdf(Scene const &s, int mat, V3<T> n, V3<T> o, V3<T> i) {
    if (mat < 0 || s.mats[mat].kind == MAT_OREN_NAYAR) {
        T c = inner(i, n);
        return c > T(0) ? c / pi_v<T>() : T(0);
    }
    Material const &m = s.mats[mat];
    if (m.kind == MAT_COOK_TORR) {
        if (inner(i, n) < T(0))
            return T(0);
        V3<T> h = i + o;
        if (near0(h))
            return T(0);
        h = normalize(h);
        T hdo = inner(h, o);
        if (!(hdo > T(0)))
            return T(0);
        T ct = inner(n, h);
        if (!(ct > T(0)))
            return T(0);
        return ggx_d(sample_alpha2_of<T>(m), n, h) * ct / (T(4.0f) * hdo);
    }
    T r = T(m.ratio);
    return r * bsdf_pdf<T>(s, m.sub0, n, o, i) + (T(1) - r) * bsdf_pdf<T>(s, m.sub1, n, o, i);
}

// Draws consumed: oren_nayar / fail-safe / cook_torr 2, mixture 1 + the chosen lobe's.
// With ref_mixture the mixture density is the chosen lobe's density times its selection
// probability (material.cpp:94-106); otherwise it is the full mixture density.
// Uniform source replaying given numbers (to feed the reference's own rng draws).
struct Replay {
    const float *u;
    uint32_t k;
    float draw() { return u[k++]; }
};

template <class T, class R>
V3<T> bsdf_sample(Scene const &s, int mat, V3<T> n, V3<T> o, R &g, T *dens, bool ref_mixture) {
    if (mat < 0 || s.mats[mat].kind == MAT_OREN_NAYAR) { // material.cpp:163-169
        T e0 = T(g.draw());
        T e1 = T(g.draw());
        V3<T> i = cos_hemisphere(n, e0, e1);
        *dens = inner(i, n) / pi_v<T>();
        return i;
    }
    Material const &m = s.mats[mat];
    if (m.kind == MAT_COOK_TORR) { // material.cpp:224-252
        T ms = sample_alpha2_of<T>(m);
        V3<T> u, v;
        basis(n, &u, &v);
        T e0 = T(g.draw());
        T t = T(g.draw());
        T st, ct;
        sincos_2pi(e0, &st, &ct);
        // phi = atan(sqrt(q)): cos(phi) = 1/sqrt(1+q), sin(phi) = sqrt(q) cos(phi).
        T q = ms * t / (T(1.0f) - t);
        T cos_phi = T(1.0f) / std::sqrt(T(1.0f) + q);
        T sin_phi = std::sqrt(q) * cos_phi;
        T x = sin_phi * ct;
        T y = sin_phi * st;
        T z = cos_phi;
        V3<T> h = x * u + y * v + z * n;
        T hdo = inner(h, o);
        V3<T> i = (T(2.0f) * hdo) * h - o;
        if (inner(i, n) < T(0.0f)) {
            *dens = T(0);
            return V3<T>();
        }
        *dens = ggx_d(ms, n, h) * cos_phi / (T(4.0f) * hdo);
        return i;
    }
    T r = T(m.ratio);
    T sel = T(g.draw());
    V3<T> i;
    T d;
    if (sel < r) {
        i = bsdf_sample<T, R>(s, m.sub0, n, o, g, &d, ref_mixture);
        d = ref_mixture ? d * r : d;
    } else {
        i = bsdf_sample<T, R>(s, m.sub1, n, o, g, &d, ref_mixture);
        d = ref_mixture ? d * (T(1) - r) : d;
    }
    if (!ref_mixture && d != T(0))
        d = bsdf_pdf<T>(s, mat, n, o, i);
    *dens = d;
    return i;
}

template V3<float> bsdf_eval<float>(Scene const &, int, V3f, V3f, V3f);
template V3<double> bsdf_eval<double>(Scene const &, int, V3<double>, V3<double>, V3<double>);
template float bsdf_pdf<float>(Scene const &, int, V3f, V3f, V3f);
template double bsdf_pdf<double>(Scene const &, int, V3<double>, V3<double>, V3<double>);
template V3<float> bsdf_sample<float, Pcg>(Scene const &, int, V3f, V3f, Pcg &, float *, bool);
template V3<double> bsdf_sample<double, Pcg>(Scene const &, int, V3<double>, V3<double>, Pcg &,
                                             double *, bool);

// ---------------------------------------------------------------------------------------------
// Lights.
// ---------------------------------------------------------------------------------------------
template <class T> struct LightSample {
    int light;
    V3<T> p, n;
    T area_dens; // area density times light selection probability (pathtracer.cpp:206-207)
    bool sky;
};

// Selection probability |power| / total power (lightsources.cpp:73); 0 when every light has zero
// power (the reference's 0 / 0).
inline float pick_prob(Scene const &s, Light const &L) {
    return s.total_power > 0.0f ? norm(L.power) / s.total_power : 0.0f;
}

// basic_light_sources::sample_light (lightsources.cpp:59-75), then a point on the light: on an
// area light trimesh::sample (geometry.cpp:44-66) with the triangle drawn by area instead of by
// index; on the sky sky_light::sample_emssion_surface (light.cpp:117-129): u cosine-distributed
// about +z, p = (u + ref) * dia, n = -u, density (z . u) / pi. Draws consumed: 4 (the sky
// skips the second).
template <class T> LightSample<T> sample_light(Scene const &s, Pcg &g) {
    LightSample<T> ls;
    float e = g.draw() * s.total_power;
    unsigned lo = 0, hi = static_cast<unsigned>(s.light_cum.size());
    while (lo < hi) {
        unsigned mi = (lo + hi) >> 1;
        if (s.light_cum[mi] < e)
            lo = mi + 1;
        else
            hi = mi;
    }
    if (lo >= s.light_cum.size())
        lo = static_cast<unsigned>(s.light_cum.size()) - 1;
    Light const &L = s.lights[lo];
    float prob = pick_prob(s, L);
    ls.light = static_cast<int>(lo);
    ls.sky = L.sky();
    if (L.sky()) {
        (void)g.draw();
        T e0 = T(g.draw());
        T e1 = T(g.draw());
        V3<T> z(T(0), T(0), T(1));
        V3<T> u = cos_hemisphere(z, e0, e1);
        ls.n = -u;
        ls.p = (u + cvt<T>(L.sky_ref)) * T(L.sky_dia);
        ls.area_dens = (inner(z, u) / pi_v<T>()) * T(prob);
        return ls;
    }
    Mesh const &m = s.meshes[L.mesh];
    unsigned nt = static_cast<unsigned>(m.cum_area.size());
    float ea = g.draw() * m.area;
    unsigned a = 0, b = nt;
    while (a < b) {
        unsigned mi = (a + b) >> 1;
        if (m.cum_area[mi] < ea)
            a = mi + 1;
        else
            b = mi;
    }
    if (a >= nt)
        a = nt - 1;
    T u = T(g.draw());
    T v = T(g.draw());
    T r = std::sqrt(u);
    T b0 = T(1) - r;
    T b1 = r * v;
    T b2 = T(1) - b0 - b1;
    uint32_t i0 = m.tris[3 * a], i1 = m.tris[3 * a + 1], i2 = m.tris[3 * a + 2];
    ls.p = b0 * cvt<T>(m.verts[i0]) + b1 * cvt<T>(m.verts[i1]) + b2 * cvt<T>(m.verts[i2]);
    ls.n = normalize(b0 * cvt<T>(m.normals[i0]) + b1 * cvt<T>(m.normals[i1]) +
                     b2 * cvt<T>(m.normals[i2]));
    ls.area_dens = (T(1.0f) / T(m.area)) * T(prob);
    return ls;
}

// area_light::eval (light.cpp:41-52): radiance through the connection l = target - p_light;
// sky_light::eval (light.cpp:131-137): no distance fall-off, cosines clamped at zero.
template <class T> V3<T> light_eval(Light const &L, V3<T> l, V3<T> nl, V3<T> nt, T *r2_out) {
    if (L.sky()) {
        V3<T> in = normalize(l);
        T cos_o = inner(nl, in);
        T cos_i = inner(nt, -in);
        *r2_out = T(0);
        return (cvt<T>(L.rad) * smax(cos_o, T(0))) * smax(cos_i, T(0));
    }
    T r2 = inner(l, l);
    V3<T> in = l / std::sqrt(r2);
    T cos_o = inner(nl, in);
    T cos_i = inner(nt, -in);
    *r2_out = r2;
    if (cos_o > 0 && cos_i > 0)
        return ((cvt<T>(L.rad) * cos_i) * cos_o) / r2;
    return V3<T>();
}

// area_light::radiance (light.cpp:62-71).
template <class T> V3<T> light_radiance(Light const &L, V3<T> w, V3<T> n) {
    return inner(n, w) > 0 ? cvt<T>(L.rad) : V3<T>();
}

template <class T> T light_pick_prob(Scene const &s, int light) { return T(pick_prob(s, s.lights[light])); }

// Emission direction of a light sample and its solid-angle density: cosine-distributed about the
// area light's normal (area_light::sample_emssion, light.cpp:26-32), or the sky's inward normal
// with density 1 (sky_light::sample_emssion, light.cpp:102-115).
template <class T> V3<T> emission_dir(LightSample<T> const &em, T e0, T e1, T *w_dens) {
    if (em.sky) {
        *w_dens = T(1);
        return em.n;
    }
    V3<T> w = cos_hemisphere(em.n, e0, e1);
    *w_dens = inner(em.n, w) / pi_v<T>();
    return w;
}

// ---------------------------------------------------------------------------------------------
// Integrators.
// ---------------------------------------------------------------------------------------------
namespace {

constexpr uint32_t kRefModeMaxDepth = 64; // RR p=0.5 from depth 2: P(depth >= 64) = 2^-62

template <class T> struct Ctx {
    Scene const &s;
    RenderOpts const &opt;
    uint64_t *stats;
};

template <class T> Hit<T> trace(Ctx<T> &c, V3<T> o, V3<T> v) {
    if (c.stats)
        c.stats[4]++;
    return intersect<T>(c.s, o, v, c.opt.traversal, c.stats);
}
template <class T> bool shadow(Ctx<T> &c, V3<T> o, V3<T> v, T tmin, T tmax) {
    if (c.stats)
        c.stats[5]++;
    return occluded<T>(c.s, o, v, tmin, tmax, c.opt.traversal == TRAV_BRUTE ? TRAV_BRUTE : TRAV_REF,
                       c.stats);
}

template <class T> int mat_of(Scene const &s, int geo) { return s.meshes[geo].material; }

// transport_illum_source (pathtracer.cpp:152-176) divided by the light sample's area density
// (pathtracer.cpp:236-239). Returns the unweighted contribution; r2/cos at the light out.
template <class T>
V3<T> direct_from(Ctx<T> &c, LightSample<T> const &ls, Hit<T> const &x, V3<T> wo, V3<T> *wi_out,
                  T *r2_out, T *cos_l_out, bool *ok) {
    *ok = false;
    V3<T> l = x.p - ls.p;
    T r2;
    V3<T> illum = light_eval<T>(c.s.lights[ls.light], l, ls.n, x.n, &r2);
    if (near0(illum))
        return V3<T>();
    T dist = norm(l);
    V3<T> wi = -l / dist;
    if (shadow<T>(c, x.p, wi, T(1e-4f), dist - T(1e-3f)))
        return V3<T>();
    V3<T> f = bsdf_eval<T>(c.s, mat_of<T>(c.s, x.geo), x.n, wo, wi);
    *ok = true;
    *wi_out = wi;
    *r2_out = r2;
    *cos_l_out = inner(ls.n, -wi);
    return mod(illum, f) / ls.area_dens;
}

// One vertex of a recorded MIS path (for the adjoint): shading normal, outgoing direction,
// sampled BSDF direction, light-sample direction, the NEE contribution made at the vertex and
// the emission reached through the BSDF sample (both already weighted by throughput and MIS).
template <class T> struct AdjVertT {
    V3<T> n, wo, wi, wl, c_nee, c_em;
    int mat = -1;
    bool nee = false, cont = false;
};
constexpr uint32_t kAdjMaxDepth = 8;

// Config mode: unidirectional path tracing with next-event estimation and BSDF sampling
// combined by the power heuristic (Veach 1997), fixed maximum of opt.max_depth bounces.
// With rec, every bounce's vertex is recorded (rec[0 .. *nrec)).
template <class T>
V3<T> li_mis(Ctx<T> &c, Hit<T> const &first, V3<T> ray_v, Pcg &g, AdjVertT<T> *rec = nullptr,
             uint32_t *nrec = nullptr) {
    Scene const &s = c.s;
    V3<T> L, beta(T(1));
    if (nrec)
        *nrec = 0;
    if (!first.valid())
        return L;
    Hit<T> x = first;
    V3<T> wo = -ray_v;
    int lx = s.meshes[x.geo].light;
    if (lx >= 0)
        L = L + light_radiance<T>(s.lights[lx], wo, x.n);
    for (uint32_t k = 0; k < c.opt.max_depth; k++) {
        int mat = mat_of<T>(s, x.geo);
        AdjVertT<T> *r = rec ? rec + k : nullptr;
        if (r) {
            *r = AdjVertT<T>();
            r->n = x.n;
            r->wo = wo;
            r->mat = mat;
            *nrec = k + 1;
        }
        if (!s.lights.empty()) {
            LightSample<T> ls = sample_light<T>(s, g);
            V3<T> wi;
            T r2 = 0, cl = 0;
            bool ok;
            V3<T> d = direct_from<T>(c, ls, x, wo, &wi, &r2, &cl, &ok);
            if (ok) {
                // a BSDF sample never reaches the sky (it has no surface): weight 1
                T w = T(1);
                if (!ls.sky) {
                    T pl = ls.area_dens * r2 / cl;
                    T pb = bsdf_pdf<T>(s, mat, x.n, wo, wi);
                    w = pl * pl / (pl * pl + pb * pb);
                }
                V3<T> add = mod(beta, d * w);
                L = L + add;
                if (r) {
                    r->nee = true;
                    r->wl = wi;
                    r->c_nee = add;
                }
            }
        }
        T dens;
        V3<T> wi = bsdf_sample<T>(s, mat, x.n, wo, g, &dens, false);
        if (!(dens > T(0)))
            break;
        Hit<T> nx = trace<T>(c, x.p, wi);
        if (!nx.valid() || inner(nx.n, -wi) <= T(0))
            break;
        T cw = inner(x.n, wi);
        V3<T> f = bsdf_eval<T>(s, mat, x.n, wo, wi);
        beta = mod(beta, f) * cw / dens;
        if (r) {
            r->cont = true;
            r->wi = wi;
        }
        int ln = s.meshes[nx.geo].light;
        if (ln >= 0) {
            T cl = inner(nx.n, -wi);
            T pl = (light_pick_prob<T>(s, ln) / T(s.meshes[s.lights[ln].mesh].area)) *
                   (nx.t * nx.t) / cl;
            T w = dens * dens / (dens * dens + pl * pl);
            V3<T> add = mod(beta, light_radiance<T>(s.lights[ln], -wi, nx.n) * w);
            L = L + add;
            if (r)
                r->c_em = add;
        }
        x = nx;
        wo = -wi;
    }
    return L;
}

// Scalar factor s of a non-mixture lobe (its value is s * albedo) and ds/droughness.
// Oren-Nayar: s = k / pi with the cut-offs of material.cpp:134-161, no roughness parameter.
// cook_torr: s = F D G / (4 cos_i cos_o) (material.cpp:171-198) with alpha^2 = u = 2 r^2:
//   dD/du = D (t - u) / (u (u + t)), t = tan^2 of the half vector's angle;
//   dG/du = -G^2 (dLambda_i + dLambda_o), dLambda/du = tan^2 / (4 sqrt(1 + u tan^2));
//   ds/dr = s (dD/du / D + dG/du / G) * 4 r.
template <class T> void lobe_scalar(Material const &m, V3<T> n, V3<T> o, V3<T> i, T *sv, T *ds_dr) {
    *sv = T(0);
    *ds_dr = T(0);
    if (m.kind == MAT_OREN_NAYAR) {
        T k;
        if (oren_nayar_k<T>(m, n, o, i, &k))
            *sv = (T(1.0f) / pi_v<T>()) * k;
        return;
    }
    T cos_o = smax(T(0.0f), inner(n, o));
    T cos_i = smax(T(0.0f), inner(n, i));
    if (near0(cos_i) || near0(cos_o))
        return;
    V3<T> h = i + o;
    if (near0(h))
        return;
    h = normalize(h);
    T u = alpha2_of<T>(m);
    T f = fresnel(m, i, h);
    T d = ggx_d(u, n, h);
    T g = ggx_g(u, i, o, n);
    *sv = f * d * g / (T(4.0f) * cos_i * cos_o);
    T ct = inner(n, h);
    T t = (T(1) - ct * ct) / (ct * ct);
    auto dlam = [&](V3<T> w) {
        T c = inner(n, w);
        T t2 = (T(1) - c * c) / (c * c);
        return t2 / (T(4) * std::sqrt(T(1) + u * t2));
    };
    T dlog = (t - u) / (u * (u + t)) - g * (dlam(i) + dlam(o));
    *ds_dr = *sv * dlog * T(4) * T(m.roughness);
}

// Adds Q-weighted derivatives of log f (per channel) of material `mat` at (n, o, i) to grad
// (4 per material: d albedo rgb, d roughness). f_c = sum over lobes w_l s_l a_{l,c}, so
// d f_c / d a_{l,c} = w_l s_l and d f_c / d r_l = w_l a_{l,c} ds_l/dr.
template <class T> void add_bsdf_grad(Scene const &s, int mat, V3<T> n, V3<T> o, V3<T> i, V3<T> Q, double *grad) {
    if (mat < 0)
        return; // fail-safe material: no parameters
    V3<T> f = bsdf_eval<T>(s, mat, n, o, i);
    T fc[3] = {f.x, f.y, f.z}, qc[3] = {Q.x, Q.y, Q.z};
    auto lobe = [&](int lm, T wgt) {
        Material const &L = s.mats[lm];
        T sv, ds;
        lobe_scalar<T>(L, n, o, i, &sv, &ds);
        T ac[3] = {T(L.albedo.x), T(L.albedo.y), T(L.albedo.z)};
        for (int c = 0; c < 3; c++) {
            if (fc[c] == T(0))
                continue;
            grad[4 * lm + c] += static_cast<double>(qc[c] * wgt * sv / fc[c]);
            if (L.kind == MAT_COOK_TORR)
                grad[4 * lm + 3] += static_cast<double>(qc[c] * wgt * ac[c] * ds / fc[c]);
        }
    };
    Material const &m = s.mats[mat];
    if (m.kind == MAT_MIXTURE) {
        lobe(m.sub0, T(m.ratio));
        lobe(m.sub1, T(1) - T(m.ratio));
    } else {
        lobe(mat, T(1));
    }
}

// Path-replay adjoint of one sample (Vicini et al. 2021, specialised to BSDF parameters with
// the sampled directions, sampling densities and MIS weights held fixed): each contribution's
// derivative is the contribution times the d log f of every BSDF evaluation it contains. The
// path is re-traced from its PCG stream; `after` sums the contributions made after the BSDF
// sample of vertex k (which all contain f at that vertex). gw = dLoss/dpixel / spp.
template <class T> void mis_adjoint(Ctx<T> &c, Hit<T> const &first, V3<T> ray_v, Pcg &g, V3<T> gw, double *grad) {
    AdjVertT<T> rec[kAdjMaxDepth];
    uint32_t n = 0;
    li_mis<T>(c, first, ray_v, g, rec, &n);
    V3<T> tail;
    for (int k = static_cast<int>(n) - 1; k >= 0; k--) {
        AdjVertT<T> const &r = rec[k];
        V3<T> after = r.c_em + tail;
        if (r.cont)
            add_bsdf_grad<T>(c.s, r.mat, r.n, r.wo, r.wi, mod(gw, after), grad);
        if (r.nee)
            add_bsdf_grad<T>(c.s, r.mat, r.n, r.wo, r.wl, mod(gw, r.c_nee), grad);
        tail = after + r.c_nee;
    }
}

// Reference mode: unidirect_lt1_path_tracer (pathtracer.cpp:587-653), NEE without MIS,
// Russian roulette p = 0.5 from depth 2, emission only at the first hit. The recursion is
// evaluated front to back: (d_k + L_{k+1} w_k) / p_k == sum_k d_k prod_{j<k} w_j / prod_{j<=k} p_j.
template <class T> V3<T> li_lt1(Ctx<T> &c, Hit<T> const &first, V3<T> ray_v, Pcg &g) {
    Scene const &s = c.s;
    V3<T> L, beta(T(1));
    if (!first.valid())
        return L;
    Hit<T> x = first;
    V3<T> wo = -ray_v;
    for (uint32_t depth = 0; depth < kRefModeMaxDepth; depth++) {
        if (depth >= 2) {
            if (g.draw() >= 0.5f)
                break;
            beta = beta / T(0.5f);
        }
        int mat = mat_of<T>(s, x.geo);
        if (!s.lights.empty()) {
            LightSample<T> ls = sample_light<T>(s, g);
            V3<T> wi;
            T r2 = 0, cl = 0;
            bool ok;
            V3<T> d = direct_from<T>(c, ls, x, wo, &wi, &r2, &cl, &ok);
            L = L + mod(beta, d);
        }
        T dens;
        V3<T> wi = bsdf_sample<T>(s, mat, x.n, wo, g, &dens, true);
        if (dens == T(0))
            break;
        Hit<T> nx = trace<T>(c, x.p, wi);
        if (!nx.valid() || inner(nx.n, -wi) <= T(0))
            break;
        T cw = inner(x.n, wi);
        V3<T> f = bsdf_eval<T>(s, mat, x.n, wo, wi);
        beta = mod(beta, f) * cw / dens;
        x = nx;
        wo = -wi;
    }
    int l0 = s.meshes[first.geo].light;
    if (l0 >= 0)
        L = L + light_radiance<T>(s.lights[l0], -ray_v, first.n);
    return L;
}

// bidirect_lt2_path_tracer (pathtracer.cpp:655-752) evaluated front to back: at each camera
// vertex join_with_light_paths (pathtracer.cpp:655-699) — one NEE sample p1 and a light path of
// one vertex (emission sample, closest hit) joined to the camera vertex, p2; the vertex's term
// is p1 + p2/2 at depth 0 and (p1 + p2)/2 deeper when the join is unoccluded and every cosine
// positive, else p1, and zero when the light path misses (the reference returns 0 there).
// sample_indirect_illum (pathtracer.cpp:701-731): Russian roulette p = 0.5 from depth 1; a
// BSDF sample hitting the scene with cos < 0 at the vertex zeroes the vertex (its term and
// the continuation); the hit is not tested for facing. A zero BSDF density ends the path with
// the vertex's term (the reference would divide by zero).
template <class T> V3<T> li_lt2(Ctx<T> &c, Hit<T> const &first, V3<T> ray_v, Pcg &g) {
    Scene const &s = c.s;
    V3<T> L, beta(T(1));
    if (!first.valid())
        return L;
    Hit<T> x = first;
    V3<T> wo = -ray_v;
    for (uint32_t depth = 0; depth < kRefModeMaxDepth; depth++) {
        if (depth >= 1) {
            if (g.draw() >= 0.5f)
                break;
            beta = beta / T(0.5f);
        }
        int mat = mat_of<T>(s, x.geo);
        V3<T> B;
        if (!s.lights.empty()) {
            LightSample<T> ls = sample_light<T>(s, g);
            V3<T> wl;
            T r2 = 0, cl = 0;
            bool ok;
            V3<T> p1 = direct_from<T>(c, ls, x, wo, &wl, &r2, &cl, &ok);
            LightSample<T> em = sample_light<T>(s, g);
            T e0 = T(g.draw());
            T e1 = T(g.draw());
            T w_dens;
            V3<T> w = emission_dir(em, e0, e1, &w_dens);
            Hit<T> lv = trace<T>(c, em.p, w);
            if (lv.valid()) {
                T cos_e = inner(em.n, w);
                V3<T> emitted = cos_e > T(0) ? cvt<T>(s.lights[em.light].rad) * cos_e : V3<T>();
                // a light of zero selection probability carries no power
                V3<T> illum = em.area_dens > T(0) ? emitted / (em.area_dens * w_dens) : V3<T>();
                V3<T> tray = -w;
                V3<T> jp = x.p - lv.p;
                T dist = norm(jp);
                jp = jp / dist;
                T cos_w2 = inner(lv.n, tray);
                T cos_wo = inner(lv.n, jp);
                T cos_wi = inner(x.n, -jp);
                if (cos_wo > T(0) && cos_wi > T(0) && cos_w2 > T(0) &&
                    !shadow<T>(c, lv.p, jp, T(1e-4f), dist - T(1e-3f))) {
                    V3<T> f2 = mod(illum, bsdf_eval<T>(s, mat_of<T>(s, lv.geo), lv.n, jp, tray)) * cos_w2;
                    V3<T> p2 = mod(f2 * cos_wo / (dist * dist), bsdf_eval<T>(s, mat, x.n, wo, -jp)) * cos_wi;
                    B = depth == 0 ? p1 + T(0.5f) * p2 : T(0.5f) * (p1 + p2);
                } else {
                    B = p1;
                }
            }
        }
        T dens;
        V3<T> wi = bsdf_sample<T>(s, mat, x.n, wo, g, &dens, true);
        if (dens == T(0)) {
            L = L + mod(beta, B);
            break;
        }
        Hit<T> nx = trace<T>(c, x.p, wi);
        if (!nx.valid()) {
            L = L + mod(beta, B);
            break;
        }
        T cw = inner(x.n, wi);
        if (cw < T(0))
            break;
        L = L + mod(beta, B);
        V3<T> f = bsdf_eval<T>(s, mat, x.n, wo, wi);
        beta = mod(beta, f) * cw / dens;
        x = nx;
        wo = -wi;
    }
    int l0 = s.meshes[first.geo].light;
    T cos0 = inner(first.n, -ray_v);
    if (l0 >= 0 && cos0 > T(0)) // area_light::projected_radiance (light.cpp:54-61)
        L = L + cvt<T>(s.lights[l0].rad) * cos0;
    return L;
}

// Reference mode: unidirect_path_tracer (pathtracer.cpp:528-585), BSDF sampling only,
// emission at every vertex, Russian roulette p = 0.5 from depth 2.
template <class T> V3<T> li_unidirect(Ctx<T> &c, Hit<T> const &first, V3<T> ray_v, Pcg &g) {
    Scene const &s = c.s;
    V3<T> L, beta(T(1));
    if (!first.valid())
        return L;
    Hit<T> x = first;
    V3<T> wo = -ray_v;
    for (uint32_t depth = 0; depth < kRefModeMaxDepth; depth++) {
        if (depth >= 2) {
            if (g.draw() >= 0.5f)
                break;
            beta = beta / T(0.5f);
        }
        int lx = s.meshes[x.geo].light;
        if (lx >= 0)
            L = L + mod(beta, light_radiance<T>(s.lights[lx], wo, x.n));
        int mat = mat_of<T>(s, x.geo);
        T dens;
        V3<T> wi = bsdf_sample<T>(s, mat, x.n, wo, g, &dens, true);
        if (dens == T(0))
            break;
        Hit<T> nx = trace<T>(c, x.p, wi);
        if (!nx.valid() || inner(nx.n, -wi) <= T(0))
            break;
        T cw = inner(x.n, wi);
        V3<T> f = bsdf_eval<T>(s, mat, x.n, wo, wi);
        beta = mod(beta, f) * cw / dens;
        x = nx;
        wo = -wi;
    }
    return L;
}

// direct_path_tracer (pathtracer.cpp:508-526): one light sample at the first hit plus the
// first hit's projected emission.
template <class T> V3<T> li_direct(Ctx<T> &c, Hit<T> const &first, V3<T> ray_v, Pcg &g) {
    Scene const &s = c.s;
    V3<T> L;
    if (!first.valid())
        return L;
    V3<T> wo = -ray_v;
    if (!s.lights.empty()) {
        LightSample<T> ls = sample_light<T>(s, g);
        V3<T> wi;
        T r2 = 0, cl = 0;
        bool ok;
        L = direct_from<T>(c, ls, first, wo, &wi, &r2, &cl, &ok);
    }
    int l0 = s.meshes[first.geo].light;
    if (l0 >= 0 && inner(first.n, wo) > T(0))
        L = L + cvt<T>(s.lights[l0].rad) * inner(first.n, wo);
    return L;
}

// ---------------------------------------------------------------------------------------------
// Bidirectional path tracing (bidirect_mis_path_tracer, pathtracer.cpp:12-137, 247-304,
// 351-452, 754-803).
// ---------------------------------------------------------------------------------------------
constexpr uint32_t kMaxPathLen = 16; // upper bound of RenderOpts::max_path_len

// sampled_pathlet (pathtracer.cpp:12-45): a subpath vertex, the unit vector from it towards the
// previous vertex (towards_prev) and the prefix transport up to it (light_transport_info,
// pathtracer.cpp:254-274), accumulated while the walk extends the subpath.
template <class T> struct Pathlet {
    Hit<T> x;
    V3<T> v;
    V3<T> transport;
};

// sample_path (pathtracer.cpp:70-92): extends path[0..len) by BSDF sampling (the reference's
// mixture densities) until max_len vertices, a density within 1e-5 of zero, a miss or a
// back-facing hit. IMPORTANCE selects projected_brdf (light subpaths) over
// projected_adjoint_brdf (camera subpaths) for the prefix transport (pathtracer.cpp:54-68,
// 259-270): transport_{k+1} = transport_k * (f * cos) / dens.
template <class T, bool IMPORTANCE>
uint32_t extend_walk(Ctx<T> &c, Pathlet<T> *path, uint32_t len, uint32_t max_len, Pcg &g) {
    Scene const &s = c.s;
    while (len < max_len) {
        Pathlet<T> const &cur = path[len - 1];
        int mat = mat_of<T>(s, cur.x.geo);
        T dens;
        V3<T> i = bsdf_sample<T>(s, mat, cur.x.n, cur.v, g, &dens, true);
        if (near0(dens))
            break;
        Hit<T> nx = trace<T>(c, cur.x.p, i);
        if (!nx.valid() || !(inner(nx.n, -i) > T(0)))
            break;
        T cos_w = inner(cur.x.n, i);
        V3<T> f = IMPORTANCE ? bsdf_eval<T>(s, mat, cur.x.n, i, cur.v) : bsdf_eval<T>(s, mat, cur.x.n, cur.v, i);
        path[len].x = nx;
        path[len].v = -i;
        path[len].transport = mod(cur.transport, (f * cos_w) / dens);
        len++;
    }
    return len;
}

// bidirect_mis_path_tracer::sample (pathtracer.cpp:775-803) for one camera ray: a camera
// subpath from the cached first hit, an emission sample (light by power, point by area,
// cosine-weighted direction; pathtracer.cpp:754-767, light.cpp:26-32) and a light subpath,
// then transport_all_connectible_subpaths (pathtracer.cpp:351-452): for each total length,
// the average over its strategies with the reference's uniform weights — (0,0) and every
// light_plen == 0 or connection strategy counts, occluded or not; cam_plen == 0 does not.
// Draws: camera walk, then 6 for the emission sample, then the light walk.
template <class T> V3<T> li_bdpt(Ctx<T> &c, Hit<T> const &first, V3<T> ray_v, Pcg &g) {
    Scene const &s = c.s;
    V3<T> rad;
    uint32_t max_len = c.opt.max_path_len;
    if (!first.valid() || max_len == 0 || s.lights.empty())
        return rad;
    Pathlet<T> cam[kMaxPathLen], lig[kMaxPathLen];
    cam[0].x = first;
    cam[0].v = -ray_v;
    cam[0].transport = V3<T>(T(1));
    uint32_t cam_len = extend_walk<T, false>(c, cam, 1, max_len, g);

    LightSample<T> em = sample_light<T>(s, g);
    T e0 = T(g.draw());
    T e1 = T(g.draw());
    T w_dens;
    V3<T> w = emission_dir(em, e0, e1, &w_dens);
    Light const &light = s.lights[em.light];
    uint32_t light_len = 0;
    Hit<T> v0 = trace<T>(c, em.p, w);
    if (v0.valid() && inner(v0.n, -w) > T(0)) {
        lig[0].x = v0;
        lig[0].v = -w;
        lig[0].transport = V3<T>(T(1));
        light_len = extend_walk<T, true>(c, lig, 1, max_len, g);
    }
    // light.projected_radiance(w, n) / (dens0 * area_dens) (pathtracer.cpp:418-420)
    T cos_e = inner(em.n, w);
    V3<T> emitted = cos_e > T(0) ? cvt<T>(light.rad) * cos_e : V3<T>();
    V3<T> light_emission = em.area_dens > T(0) ? emitted / (w_dens * em.area_dens) : V3<T>();

    for (uint32_t plen = 1; plen <= cam_len + light_len + 1; plen++) {
        int cp = static_cast<int>(std::min(plen - 1, cam_len));
        uint32_t lp = plen - 1 - static_cast<uint32_t>(cp);
        V3<T> psum;
        T wsum = T(0);
        for (; cp >= 0 && lp <= light_len; lp++, cp--) {
            if (lp == 0 && cp == 0) {
                int l0 = s.meshes[cam[0].x.geo].light;
                if (l0 >= 0)
                    psum = psum + light_radiance<T>(s.lights[l0], cam[0].v, cam[0].x.n);
                wsum = wsum + T(1);
            } else if (lp == 0) {
                Pathlet<T> const &cj = cam[cp - 1];
                V3<T> wi;
                T r2 = 0, cl = 0;
                bool ok;
                V3<T> tr = direct_from<T>(c, em, cj.x, cj.v, &wi, &r2, &cl, &ok);
                psum = psum + mod(tr, cj.transport);
                wsum = wsum + T(1);
            } else if (cp > 0) {
                Pathlet<T> const &lj = lig[lp - 1];
                Pathlet<T> const &cj = cam[cp - 1];
                V3<T> join = cj.x.p - lj.x.p;
                T dist = norm(join);
                join = join / dist;
                T cos_wo = inner(lj.x.n, join);
                T cos_wi = inner(cj.x.n, -join);
                if (cos_wo > T(0) && cos_wi > T(0) && !shadow<T>(c, lj.x.p, join, T(1e-3f), dist - T(1e-3f))) {
                    V3<T> sub = mod(light_emission, lj.transport);
                    T geo = cos_wi * cos_wo / (dist * dist);
                    V3<T> fl = bsdf_eval<T>(s, mat_of<T>(s, lj.x.geo), lj.x.n, join, lj.v);
                    V3<T> fc = bsdf_eval<T>(s, mat_of<T>(s, cj.x.geo), cj.x.n, cj.v, -join);
                    V3<T> tr = mod(mod(sub, fl), fc) * geo;
                    psum = psum + mod(tr, cj.transport);
                }
                wsum = wsum + T(1);
            }
        }
        if (wsum > T(0))
            rad = rad + psum / wsum;
    }
    return rad;
}

} // namespace

template <class T>
V3<T> trace_sample(Scene const &s, RenderOpts const &opt, Hit<T> const &first, V3<T> const &ray_v,
                   uint64_t pixel, uint32_t sample, uint64_t *stats) {
    Ctx<T> c{s, opt, stats};
    Pcg g = pcg_for(opt.seed, pixel, sample);
    switch (opt.tracer) {
    case TR_UNIDIRECT_MIS:
        return li_mis<T>(c, first, ray_v, g);
    case TR_UNIDIRECT_LT1:
        return li_lt1<T>(c, first, ray_v, g);
    case TR_UNIDIRECT:
        return li_unidirect<T>(c, first, ray_v, g);
    case TR_DIRECT:
        return li_direct<T>(c, first, ray_v, g);
    case TR_BIDIRECT_MIS:
        return li_bdpt<T>(c, first, ray_v, g);
    case TR_BIDIRECT_LT2:
        return li_lt2<T>(c, first, ray_v, g);
    case TR_POSITION: { // position_tracer (pathtracer.cpp:474-490): (p - min) / (max - min)
        if (!first.valid())
            return V3<T>();
        V3<T> lo = cvt<T>(s.space_lo), r = cvt<T>(s.space_range);
        return V3<T>((first.p.x - lo.x) / r.x, (first.p.y - lo.y) / r.y, (first.p.z - lo.z) / r.z);
    }
    case TR_NORMAL: // normal_tracer (pathtracer.cpp:492-506)
        return first.valid() ? (first.n + V3<T>(T(1.0f))) / T(2.0f) : V3<T>();
    default:
        return V3<T>();
    }
}

template V3<float> trace_sample<float>(Scene const &, RenderOpts const &, Hit<float> const &,
                                       V3f const &, uint64_t, uint32_t, uint64_t *);
template V3<double> trace_sample<double>(Scene const &, RenderOpts const &, Hit<double> const &,
                                         V3<double> const &, uint64_t, uint32_t, uint64_t *);

} // namespace orc

// =============================================================================================
// C API (ctypes), mirroring oracle/ref_harness.cpp where the two overlap.
// =============================================================================================
using namespace orc;

namespace {
V3f ld3(float const *p) { return V3f(p[0], p[1], p[2]); }
template <class T> void st3(float *d, V3<T> v) {
    d[0] = static_cast<float>(v.x);
    d[1] = static_cast<float>(v.y);
    d[2] = static_cast<float>(v.z);
}
template <class T> void st3d(double *d, V3<T> v) {
    d[0] = v.x;
    d[1] = v.y;
    d[2] = v.z;
}
} // namespace

extern "C" {

void *orc_scene_new() { return new Scene(); }
void orc_scene_free(void *h) { delete static_cast<Scene *>(h); }

int orc_add_oren_nayar(void *h, float const *albedo, float sigma) {
    Scene *s = static_cast<Scene *>(h);
    Material m;
    m.kind = MAT_OREN_NAYAR;
    m.albedo = ld3(albedo);
    m.sigma = sigma;
    s->mats.push_back(m);
    return static_cast<int>(s->mats.size()) - 1;
}
int orc_add_cook_torr(void *h, float const *albedo, float roughness, float ior_re, float ior_im) {
    Scene *s = static_cast<Scene *>(h);
    Material m;
    m.kind = MAT_COOK_TORR;
    m.albedo = ld3(albedo);
    m.roughness = roughness;
    m.sample_roughness = roughness;
    m.ior_re = ior_re;
    m.ior_im = ior_im;
    s->mats.push_back(m);
    return static_cast<int>(s->mats.size()) - 1;
}
// x3d_scene_set_material_params: albedo (Oren-Nayar, cook_torr) and roughness (cook_torr) of
// materials [first, first + count); update_sampling also moves the GGX sampling roughness.
void orc_set_material_params(void *h, uint32_t first, uint32_t count, float const *albedo, float const *roughness,
                             int update_sampling) {
    Scene *s = static_cast<Scene *>(h);
    for (uint32_t k = 0; k < count; k++) {
        Material &m = s->mats[first + k];
        if (m.kind == MAT_MIXTURE)
            continue;
        if (albedo)
            m.albedo = ld3(albedo + 3 * k);
        if (roughness && m.kind == MAT_COOK_TORR) {
            m.roughness = roughness[k];
            if (update_sampling)
                m.sample_roughness = roughness[k];
        }
    }
}

int orc_add_mixture(void *h, int m0, int m1, float ratio) {
    Scene *s = static_cast<Scene *>(h);
    Material m;
    m.kind = MAT_MIXTURE;
    m.sub0 = m0;
    m.sub1 = m1;
    m.ratio = ratio;
    s->mats.push_back(m);
    return static_cast<int>(s->mats.size()) - 1;
}
int orc_add_mesh(void *h, float const *verts, float const *normals, uint32_t nv,
                 uint32_t const *tris, uint32_t nt, int mat) {
    Scene *s = static_cast<Scene *>(h);
    Mesh m;
    // The reference loads geometry through an identity transform (pathspace.cpp:15-19,
    // geometry.cpp:76-87), which turns -0 into +0; adding +0 does the same here.
    for (uint32_t i = 0; i < nv; i++) {
        m.verts.push_back(V3f(verts[3 * i] + 0.0f, verts[3 * i + 1] + 0.0f, verts[3 * i + 2] + 0.0f));
        m.normals.push_back(
            V3f(normals[3 * i] + 0.0f, normals[3 * i + 1] + 0.0f, normals[3 * i + 2] + 0.0f));
    }
    m.tris.assign(tris, tris + 3 * nt);
    m.material = mat;
    s->meshes.push_back(std::move(m));
    return static_cast<int>(s->meshes.size()) - 1;
}
int orc_add_area_light(void *h, int mesh, float const *rad) {
    Scene *s = static_cast<Scene *>(h);
    s->lights.push_back({mesh, ld3(rad), V3f()});
    return static_cast<int>(s->lights.size()) - 1;
}
// sky_light (light.cpp:87-163), placed at commit by set_scene_boundary with the path space's
// bounds.
int orc_add_sky_light(void *h, float const *rad) {
    Scene *s = static_cast<Scene *>(h);
    s->lights.push_back({-1, ld3(rad), V3f()});
    return static_cast<int>(s->lights.size()) - 1;
}
// Light state after commit: power (3), sky reference point (3), sky diameter.
void orc_light_info(void *h, int light, float *out) {
    Light const &L = static_cast<Scene *>(h)->lights.at(static_cast<size_t>(light));
    float v[7] = {L.power.x, L.power.y, L.power.z, L.sky_ref.x, L.sky_ref.y, L.sky_ref.z, L.sky_dia};
    std::memcpy(out, v, sizeof v);
}
// Light evaluation for tests: light_eval of light `light` at connections l (float[3n]), light
// normals nl and target normals nt -> out float[3n].
void orc_light_eval(void *h, int light, uint32_t n, float const *l, float const *nl, float const *nt, float *out) {
    Light const &L = static_cast<Scene *>(h)->lights.at(static_cast<size_t>(light));
    for (uint32_t k = 0; k < n; k++) {
        float r2;
        V3f v = light_eval<float>(L, ld3(l + 3 * k), ld3(nl + 3 * k), ld3(nt + 3 * k), &r2);
        out[3 * k] = v.x;
        out[3 * k + 1] = v.y;
        out[3 * k + 2] = v.z;
    }
}
int orc_commit(void *h) {
    commit(*static_cast<Scene *>(h));
    return 0;
}
void orc_bvh_size(void *h, uint32_t *nn, uint32_t *np) {
    Scene *s = static_cast<Scene *>(h);
    *nn = static_cast<uint32_t>(s->nodes.size());
    *np = static_cast<uint32_t>(s->prims.size());
}
void orc_bvh_dump(void *h, float *bounds, uint32_t *meta, int32_t *prims) {
    Scene *s = static_cast<Scene *>(h);
    for (size_t n = 0; n < s->nodes.size(); n++) {
        st3(bounds + 6 * n, s->nodes[n].bound.mn);
        st3(bounds + 6 * n + 3, s->nodes[n].bound.mx);
        meta[4 * n + 0] = s->nodes[n].num_prims;
        meta[4 * n + 1] = s->nodes[n].split_axis;
        meta[4 * n + 2] = s->nodes[n].next_child;
        meta[4 * n + 3] = s->nodes[n].prim_start;
    }
    for (size_t i = 0; i < s->prims.size(); i++) {
        prims[2 * i] = static_cast<int32_t>(s->prims[i].geo);
        prims[2 * i + 1] = static_cast<int32_t>(s->prims[i].tri);
    }
}

// Closest hit. prim = index into the BVH primitive array (-1 miss); stats (optional):
// uint64[6] += {interior pops, triangle tests, -, -, closest queries, any-hit queries}.
void orc_intersect(void *h, float const *rays, uint32_t n, int mode, int32_t *geo, int32_t *tri,
                   int32_t *prim, float *t, float *b, float *vertex, float *normal,
                   uint64_t *stats) {
    Scene *s = static_cast<Scene *>(h);
    for (uint32_t k = 0; k < n; k++) {
        Hit<float> hh = intersect<float>(*s, ld3(rays + 6 * k), ld3(rays + 6 * k + 3), mode, stats);
        prim[k] = hh.prim;
        if (hh.valid()) {
            geo[k] = static_cast<int32_t>(s->prims[hh.prim].geo);
            tri[k] = static_cast<int32_t>(s->prims[hh.prim].tri);
            t[k] = hh.t;
            b[2 * k] = hh.b1;
            b[2 * k + 1] = hh.b2;
            st3(vertex + 3 * k, hh.p);
            st3(normal + 3 * k, hh.n);
        } else {
            geo[k] = tri[k] = -1;
            t[k] = INFINITY;
            b[2 * k] = b[2 * k + 1] = 0;
            st3(vertex + 3 * k, V3f());
            st3(normal + 3 * k, V3f());
        }
    }
}

void orc_occluded(void *h, float const *rays, float const *tmin, float const *tmax, uint32_t n,
                  int mode, uint8_t *out, uint64_t *stats) {
    Scene *s = static_cast<Scene *>(h);
    for (uint32_t k = 0; k < n; k++)
        out[k] = occluded<float>(*s, ld3(rays + 6 * k), ld3(rays + 6 * k + 3), tmin[k], tmax[k],
                                 mode, stats)
                     ? 1
                     : 0;
}

// BSDF evaluation in FP32 (prec 0) or FP64 (prec 1); mat = -1 is the fail-safe material.
void orc_bsdf_eval(void *h, int mat, uint32_t n, float const *nrm, float const *o, float const *i,
                   int prec, double *f) {
    Scene *s = static_cast<Scene *>(h);
    for (uint32_t k = 0; k < n; k++) {
        if (prec == 0)
            st3d(f + 3 * k, bsdf_eval<float>(*s, mat, ld3(nrm + 3 * k), ld3(o + 3 * k), ld3(i + 3 * k)));
        else
            st3d(f + 3 * k, bsdf_eval<double>(*s, mat, cvt<double>(ld3(nrm + 3 * k)),
                                              cvt<double>(ld3(o + 3 * k)), cvt<double>(ld3(i + 3 * k))));
    }
}
void orc_bsdf_pdf(void *h, int mat, uint32_t n, float const *nrm, float const *o, float const *i,
                  int prec, double *p) {
    Scene *s = static_cast<Scene *>(h);
    for (uint32_t k = 0; k < n; k++) {
        if (prec == 0)
            p[k] = bsdf_pdf<float>(*s, mat, ld3(nrm + 3 * k), ld3(o + 3 * k), ld3(i + 3 * k));
        else
            p[k] = bsdf_pdf<double>(*s, mat, cvt<double>(ld3(nrm + 3 * k)),
                                    cvt<double>(ld3(o + 3 * k)), cvt<double>(ld3(i + 3 * k)));
    }
}
// Sample k uses the PCG stream (seed, k, 0). ref_mixture selects the reference's mixture density.
void orc_bsdf_sample(void *h, int mat, uint64_t seed, uint32_t n, float const *nrm, float const *o,
                     int prec, int ref_mixture, double *i_out, double *dens) {
    Scene *s = static_cast<Scene *>(h);
    for (uint32_t k = 0; k < n; k++) {
        Pcg g = pcg_for(seed, k, 0);
        if (prec == 0) {
            float d;
            V3f r = bsdf_sample<float>(*s, mat, ld3(nrm + 3 * k), ld3(o + 3 * k), g, &d, ref_mixture);
            st3d(i_out + 3 * k, r);
            dens[k] = d;
        } else {
            double d;
            V3<double> r = bsdf_sample<double>(*s, mat, cvt<double>(ld3(nrm + 3 * k)),
                                               cvt<double>(ld3(o + 3 * k)), g, &d, ref_mixture);
            st3d(i_out + 3 * k, r);
            dens[k] = d;
        }
    }
}

// Like orc_bsdf_sample, but sample k consumes the uniforms u[per_sample*k ...] in order.
void orc_bsdf_sample_u(void *h, int mat, uint32_t n, float const *nrm, float const *o,
                       float const *u, uint32_t per_sample, int ref_mixture, double *i_out,
                       double *dens) {
    Scene *s = static_cast<Scene *>(h);
    for (uint32_t k = 0; k < n; k++) {
        Replay g{u + static_cast<size_t>(per_sample) * k, 0};
        float d;
        V3f r = bsdf_sample<float, Replay>(*s, mat, ld3(nrm + 3 * k), ld3(o + 3 * k), g, &d,
                                           ref_mixture);
        st3d(i_out + 3 * k, r);
        dens[k] = d;
    }
}

void orc_pcg_draws(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t n, uint32_t *out) {
    Pcg g = pcg_for(seed, pixel, sample);
    for (uint32_t k = 0; k < n; k++)
        out[k] = g.next();
}

void orc_sincos_2pi(uint32_t n, float const *u, float *s, float *c) {
    for (uint32_t k = 0; k < n; k++)
        sincos_2pi(u[k], s + k, c + k);
}

void orc_camera_rays(float sensor, float focal, float aspect, float const *trans, uint32_t w,
                     uint32_t h, float *rays) {
    Camera c{sensor, focal, aspect, {}};
    std::memcpy(c.m, trans, sizeof(c.m));
    for (uint32_t j = 0; j < h; j++)
        for (uint32_t i = 0; i < w; i++) {
            V3f o, v;
            camera_ray(c, i, j, w, h, &o, &v);
            st3(rays + 6 * (i + j * w), o);
            st3(rays + 6 * (i + j * w) + 3, v);
        }
}

// Renders the listed pixels (indices i + j*w; all pixels when pixels == nullptr) with `spp`
// samples each, accumulating sequentially per pixel and scaling by 1/spp (renderer.cpp:43-45,
// 140-159 with one task). rgba: float[4*npix] (alpha = 1/spp as in renderer.cpp:148, 154-157),
// or double when prec == 1. Work is spread over `threads` CPU threads by pixel; results do not
// depend on the thread count. stats: uint64[6] (see orc_intersect), may be null.
// if_path_tracer::compute_first_hit + sample (pathtracer.h:44-63, pathtracer.cpp:456-472) on
// caller rays (float[6*n]): samples [s_begin, s_begin + num) of every ray, sample s of ray k on
// PCG stream (seed, k, s). out: float[3 * n * num], sample-major. FP32 only.
void orc_trace_rays(void *h, float const *rays, uint32_t n, int tracer, uint32_t max_depth, uint32_t max_path_len,
                    uint64_t seed, uint32_t s_begin, uint32_t num, uint32_t threads, float *out) {
    Scene *s = static_cast<Scene *>(h);
    RenderOpts opt;
    opt.tracer = tracer;
    opt.max_depth = max_depth;
    opt.max_path_len = max_path_len;
    opt.seed = seed;
    opt.traversal = TRAV_CULLED;
    if (threads == 0)
        threads = 1;
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (;;) {
            uint32_t k = next.fetch_add(1);
            if (k >= n)
                break;
            V3f o = ld3(rays + 6 * k), v = ld3(rays + 6 * k + 3);
            Hit<float> first = intersect<float>(*s, o, v, TRAV_CULLED, nullptr);
            if (inner(first.n, -v) <= 0.0f)
                first = Hit<float>();
            for (uint32_t j = 0; j < num; j++)
                st3(out + 3 * (static_cast<size_t>(j) * n + k),
                    trace_sample<float>(*s, opt, first, v, k, s_begin + j, nullptr));
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; t++)
        pool.emplace_back(work);
    for (auto &t : pool)
        t.join();
}

void orc_render(void *h, float const *cam, int tracer, uint32_t max_depth, uint32_t max_path_len,
                uint64_t seed, int traversal, uint32_t w, uint32_t hh, uint32_t spp,
                uint32_t const *pixels, uint32_t npix, int prec, uint32_t threads, void *rgba,
                uint64_t *stats) {
    Scene *s = static_cast<Scene *>(h);
    Camera c{cam[0], cam[1], cam[2], {}};
    std::memcpy(c.m, cam + 3, sizeof(c.m));
    RenderOpts opt;
    opt.tracer = tracer;
    opt.max_depth = max_depth;
    opt.max_path_len = max_path_len;
    opt.seed = seed;
    opt.traversal = traversal;
    uint32_t total = pixels ? npix : w * hh;
    if (threads == 0)
        threads = 1;
    std::atomic<uint32_t> next{0};
    std::vector<std::array<uint64_t, 6>> tstats(threads);
    auto work = [&](uint32_t tid) {
        uint64_t *st = stats ? tstats[tid].data() : nullptr;
        if (st)
            for (auto &x : tstats[tid])
                x = 0;
        for (;;) {
            uint32_t q = next.fetch_add(1);
            if (q >= total)
                break;
            uint32_t pix = pixels ? pixels[q] : q;
            uint32_t x = pix % w, y = pix / w;
            V3f o, v;
            camera_ray(c, x, y, w, hh, &o, &v);
            float scale = 1.0f / spp;
            if (prec == 0) {
                // compute_first_hit (pathtracer.cpp:456-472)
                Hit<float> first = intersect<float>(*s, o, v, traversal, st);
                if (st)
                    st[4]++;
                if (inner(first.n, -v) <= 0.0f)
                    first = Hit<float>();
                V3f acc;
                for (uint32_t k = 0; k < spp; k++)
                    acc = acc + trace_sample<float>(*s, opt, first, v, pix, k, st);
                float *out = static_cast<float *>(rgba) + 4 * q;
                st3(out, acc * scale);
                out[3] = scale;
            } else {
                V3<double> od = cvt<double>(o), vd = cvt<double>(v);
                Hit<double> first = intersect<double>(*s, od, vd, traversal, st);
                if (st)
                    st[4]++;
                if (inner(first.n, -vd) <= 0.0)
                    first = Hit<double>();
                V3<double> acc;
                for (uint32_t k = 0; k < spp; k++)
                    acc = acc + trace_sample<double>(*s, opt, first, vd, pix, k, st);
                double *out = static_cast<double *>(rgba) + 4 * q;
                st3d(out, acc / double(spp));
                out[3] = 1.0 / spp;
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; t++)
        pool.emplace_back(work, t);
    for (auto &t : pool)
        t.join();
    if (stats)
        for (auto &ts : tstats)
            for (int i = 0; i < 6; i++)
                stats[i] += ts[i];
}

// Firefly filter of sampling_task::run (renderer.cpp:46-83) over sample images followed by the
// per-pixel in-order sum: samples = float[spp][h][w][3]; sum = float[h][w][3]. For an interior
// pixel, the sample is capped component by component at the mean of its 8 neighbours' squared
// norms plus FireFlyMinDiff = 1.5 (renderer.h:118) when its own squared norm exceeds every
// neighbour's by more than 1.5; edge pixels are added unfiltered.
void orc_firefly_accumulate(float const *samples, uint32_t w, uint32_t h, uint32_t spp, float *sum) {
    const float kMinDiff = 1.5f;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            V3f acc;
            for (uint32_t k = 0; k < spp; k++) {
                float const *img = samples + static_cast<size_t>(k) * w * h * 3;
                auto at = [&](uint32_t i, uint32_t j) { return ld3(img + 3 * (i + j * w)); };
                V3f c = at(x, y);
                if (x > 0 && x < w - 1 && y > 0 && y < h - 1) {
                    float r00 = inner(at(x - 1, y - 1), at(x - 1, y - 1)), r10 = inner(at(x, y - 1), at(x, y - 1));
                    float r20 = inner(at(x + 1, y - 1), at(x + 1, y - 1)), r01 = inner(at(x - 1, y), at(x - 1, y));
                    float r11 = inner(c, c), r21 = inner(at(x + 1, y), at(x + 1, y));
                    float r02 = inner(at(x - 1, y + 1), at(x - 1, y + 1)), r12 = inner(at(x, y + 1), at(x, y + 1));
                    float r22 = inner(at(x + 1, y + 1), at(x + 1, y + 1));
                    bool firefly = (r11 - r00) > kMinDiff && (r11 - r10) > kMinDiff && (r11 - r20) > kMinDiff &&
                                   (r11 - r01) > kMinDiff && (r11 - r21) > kMinDiff && (r11 - r02) > kMinDiff &&
                                   (r11 - r12) > kMinDiff && (r11 - r22) > kMinDiff;
                    if (firefly) {
                        float cap = 1.0f / 8.0f * (r00 + r10 + r20 + r01 + r21 + r02 + r12 + r22) + kMinDiff;
                        c = V3f(cap < c.x ? cap : c.x, cap < c.y ? cap : c.y, cap < c.z ? cap : c.z);
                    }
                }
                acc = acc + c;
            }
            st3(sum + 3 * (x + y * w), acc);
        }
}

// orc_render (FP32 paths, all pixels) with the firefly filter: every sample image is rendered,
// then filtered and summed per pixel in sample order (orc_firefly_accumulate).
void orc_render_firefly(void *h, float const *cam, int tracer, uint32_t max_depth, uint32_t max_path_len,
                        uint64_t seed, uint32_t w, uint32_t hh, uint32_t spp, uint32_t threads, float *rgba) {
    Scene *s = static_cast<Scene *>(h);
    Camera c{cam[0], cam[1], cam[2], {}};
    std::memcpy(c.m, cam + 3, sizeof(c.m));
    RenderOpts opt;
    opt.tracer = tracer;
    opt.max_depth = max_depth;
    opt.max_path_len = max_path_len;
    opt.seed = seed;
    opt.traversal = TRAV_CULLED;
    const uint32_t total = w * hh;
    std::vector<float> samples(static_cast<size_t>(spp) * total * 3);
    if (threads == 0)
        threads = 1;
    std::atomic<uint32_t> next{0};
    auto work = [&](uint32_t) {
        for (;;) {
            uint32_t pix = next.fetch_add(1);
            if (pix >= total)
                break;
            V3f o, v;
            camera_ray(c, pix % w, pix / w, w, hh, &o, &v);
            Hit<float> first = intersect<float>(*s, o, v, TRAV_CULLED, nullptr);
            if (inner(first.n, -v) <= 0.0f)
                first = Hit<float>();
            for (uint32_t k = 0; k < spp; k++)
                st3(samples.data() + 3 * (static_cast<size_t>(k) * total + pix),
                    trace_sample<float>(*s, opt, first, v, pix, k, nullptr));
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; t++)
        pool.emplace_back(work, t);
    for (auto &t : pool)
        t.join();
    std::vector<float> sum(3 * static_cast<size_t>(total));
    orc_firefly_accumulate(samples.data(), w, hh, spp, sum.data());
    const float scale = 1.0f / spp;
    for (uint32_t pix = 0; pix < total; pix++) {
        V3f a = ld3(sum.data() + 3 * pix);
        st3(rgba + 4 * pix, a * scale);
        rgba[4 * pix + 3] = scale;
    }
}

// Gradient of a film loss with respect to material parameters for the config-mode MIS tracer
// (x3d_render_backward): dl = dLoss/dfilm, float[4*w*h] (rgb used); grad = double[4*nmat]
// (d albedo rgb, d roughness per material), overwritten. prec 0 runs the paths in FP32, 1 in
// FP64; the sums are double either way and do not depend on the thread count.
void orc_render_backward(void *h, float const *cam, uint32_t max_depth, uint64_t seed, uint32_t w, uint32_t hh,
                         uint32_t spp, float const *dl, int prec, uint32_t threads, double *grad) {
    Scene *s = static_cast<Scene *>(h);
    Camera c{cam[0], cam[1], cam[2], {}};
    std::memcpy(c.m, cam + 3, sizeof(c.m));
    RenderOpts opt;
    opt.tracer = TR_UNIDIRECT_MIS;
    opt.max_depth = max_depth;
    opt.seed = seed;
    uint32_t nm = static_cast<uint32_t>(s->mats.size());
    uint32_t total = w * hh;
    if (threads == 0)
        threads = 1;
    // pixel rows are split into fixed per-thread ranges so the double sums are reproducible
    std::vector<std::vector<double>> part(threads, std::vector<double>(4 * nm, 0.0));
    auto work = [&](uint32_t tid) {
        double *g = part[tid].data();
        uint32_t begin = static_cast<uint32_t>(static_cast<uint64_t>(total) * tid / threads);
        uint32_t end = static_cast<uint32_t>(static_cast<uint64_t>(total) * (tid + 1) / threads);
        for (uint32_t pix = begin; pix < end; pix++) {
            V3f o, v;
            camera_ray(c, pix % w, pix / w, w, hh, &o, &v);
            float inv = 1.0f / static_cast<float>(spp);
            if (prec == 0) {
                Ctx<float> cx{*s, opt, nullptr};
                Hit<float> first = intersect<float>(*s, o, v, TRAV_CULLED, nullptr);
                if (inner(first.n, -v) <= 0.0f)
                    first = Hit<float>();
                V3f gw(dl[4 * pix] * inv, dl[4 * pix + 1] * inv, dl[4 * pix + 2] * inv);
                for (uint32_t k = 0; k < spp; k++) {
                    Pcg pg = pcg_for(seed, pix, k);
                    mis_adjoint<float>(cx, first, v, pg, gw, g);
                }
            } else {
                Ctx<double> cx{*s, opt, nullptr};
                V3<double> od = cvt<double>(o), vd = cvt<double>(v);
                Hit<double> first = intersect<double>(*s, od, vd, TRAV_CULLED, nullptr);
                if (inner(first.n, -vd) <= 0.0)
                    first = Hit<double>();
                double sd = static_cast<double>(spp);
                V3<double> gw(static_cast<double>(dl[4 * pix]) / sd, static_cast<double>(dl[4 * pix + 1]) / sd,
                              static_cast<double>(dl[4 * pix + 2]) / sd);
                for (uint32_t k = 0; k < spp; k++) {
                    Pcg pg = pcg_for(seed, pix, k);
                    mis_adjoint<double>(cx, first, vd, pg, gw, g);
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; t++)
        pool.emplace_back(work, t);
    for (auto &t : pool)
        t.join();
    for (uint32_t i = 0; i < 4 * nm; i++) {
        double acc = 0.0;
        for (uint32_t t = 0; t < threads; t++)
            acc += part[t][i];
        grad[i] = acc;
    }
}

} // extern "C"

extern "C" {
// pcg32_srandom_r(initstate, initseq) followed by n pcg32_random_r draws (for pinning the
// generator against its published reference outputs).
void orc_pcg_raw(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t *out) {
    orc::Pcg g;
    g.state = 0;
    g.inc = (initseq << 1u) | 1u;
    g.next();
    g.state += initstate;
    g.next();
    for (uint32_t k = 0; k < n; k++)
        out[k] = g.next();
}
} // extern "C"
