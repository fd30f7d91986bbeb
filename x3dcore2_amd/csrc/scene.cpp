// Host-side scene commit: SAH BVH build reproducing the reference layout, light tables and the
// HBM packing the kernels read (scene_layout.h).
//
// BVH build, following bvh_path_space_layout::commit (corelib/src/pathspace.cpp:121-328):
// 20 SAH buckets on the widest axis, cost 1 + 8 * sum(SA_part / SA * count), split when cheaper
// than 8 * n or when n > 4, std::partition at the chosen bucket boundary, or std::sort on the
// centroid and a median cut when depth > log2(N). The flattened pre-order array (left child at
// n + 1, right child at next_child) and the primitive permutation come out identical to the
// reference's, so primitive indices and hit ids can be compared directly. The construction is
// written directly in pre-order, and the bucket costs are computed with one prefix and one
// suffix sweep instead of the reference's quadratic loop; both produce the same floats.

#include "scene.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <functional>
#include <cstring>

namespace x3d {

namespace {

// aabb (corelib/src/tensor.h:733-813) with its default-constructed empty box (min = max = 0).
struct Box {
    float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    bool empty = true;
};
inline float fmin_std(float a, float b) { return (b < a) ? b : a; } // std::min
inline float fmax_std(float a, float b) { return (a < b) ? b : a; } // std::max
inline Box make_box(const float *a, const float *b) {
    Box r;
    for (int i = 0; i < 3; i++) {
        r.mn[i] = a[i];
        r.mx[i] = b[i];
    }
    r.empty = a[0] > b[0] || a[1] > b[1] || a[2] > b[2];
    return r;
}
// aabb::operator+(aabb): a non-empty box takes min/max with the other's stored corners even
// when the other box is empty.
inline Box join(Box const &a, Box const &b) {
    if (a.empty)
        return b;
    float lo[3], hi[3];
    for (int i = 0; i < 3; i++) {
        lo[i] = fmin_std(a.mn[i], b.mn[i]);
        hi[i] = fmax_std(a.mx[i], b.mx[i]);
    }
    return make_box(lo, hi);
}
inline Box join_point(Box const &a, HostVec3 p) {
    float q[3] = {p.x, p.y, p.z};
    if (a.empty)
        return make_box(q, q);
    float lo[3], hi[3];
    for (int i = 0; i < 3; i++) {
        lo[i] = fmin_std(a.mn[i], q[i]);
        hi[i] = fmax_std(a.mx[i], q[i]);
    }
    return make_box(lo, hi);
}
inline float area_of(Box const &b) { // aabb::surf_area
    if (b.empty)
        return 0.0f;
    float d0 = b.mx[0] - b.mn[0], d1 = b.mx[1] - b.mn[1], d2 = b.mx[2] - b.mn[2];
    return 2.0f * (d0 * d1 + d0 * d2 + d1 * d2);
}

struct BuildPrim {
    PrimRef ref;
    Box bound;
    float c[3]; // centroid (max + min) / 2
};

constexpr unsigned kBuckets = 20, kLeafMax = 4, kTriCost = 8, kBoxCost = 1;

struct Builder {
    std::vector<BuildPrim> &p;
    std::vector<FlatNode> &out;
    double log2n;
    uint32_t max_depth = 0;

    void emit_leaf(Box const &b, unsigned start, unsigned count, unsigned depth) {
        FlatNode f;
        std::memcpy(f.mn, b.mn, sizeof f.mn);
        std::memcpy(f.mx, b.mx, sizeof f.mx);
        f.num_prims = count;
        f.split_axis = 0xFF;
        f.next_child = 0xFFFFFFFFu;
        f.prim_start = start;
        out.push_back(f);
        max_depth = std::max(max_depth, depth + 1);
    }

    // Costs of the 19 bucket boundaries (pathspace.cpp:147-164) with prefix / suffix sweeps.
    // For the suffix, the reference's left fold over buckets i+1..19 equals the union of the
    // non-empty buckets, plus the origin when an empty bucket follows the first non-empty one.
    void bucket_costs(Box const *bk_box, unsigned const *bk_n, Box const &b, float *cost) {
        float sab = area_of(b);
        Box first;
        unsigned cf = 0;
        float first_area[kBuckets];
        unsigned first_n[kBuckets];
        for (unsigned i = 0; i < kBuckets - 1; i++) {
            first = join(first, bk_box[i]);
            cf += bk_n[i];
            first_area[i] = area_of(first);
            first_n[i] = cf;
        }
        Box uni;
        bool any_nonempty = false, any_empty = false, origin = false;
        unsigned cs = 0;
        for (int i = kBuckets - 2; i >= 0; i--) {
            unsigned j = static_cast<unsigned>(i) + 1; // bucket prepended to the suffix
            cs += bk_n[j];
            if (!bk_box[j].empty) {
                origin = any_empty;
                uni = any_nonempty ? join(bk_box[j], uni) : bk_box[j];
                any_nonempty = true;
            }
            any_empty = any_empty || bk_box[j].empty;
            Box second;
            if (any_nonempty) {
                second = uni;
                if (origin) {
                    HostVec3 o{0.0f, 0.0f, 0.0f};
                    second = join_point(second, o);
                }
            }
            cost[i] = kBoxCost + kTriCost * (+first_area[i] / sab * first_n[i] +
                                             area_of(second) / sab * cs);
        }
        cost[kBuckets - 1] = INFINITY;
    }

    void build(unsigned start, unsigned end, unsigned depth) {
        unsigned n = end - start;
        if (n == 1) {
            emit_leaf(p[start].bound, start, 1, depth);
            return;
        }
        Box b;
        for (unsigned i = start; i < end; i++)
            b = join(b, p[i].bound);
        float range[3] = {b.mx[0] - b.mn[0], b.mx[1] - b.mn[1], b.mx[2] - b.mn[2]};
        unsigned axis = (range[0] > range[1] && range[0] > range[2]) ? 0 : (range[1] > range[2] ? 1 : 2);
        float cost_leaf = n * kTriCost;

        Box bk_box[kBuckets];
        unsigned bk_n[kBuckets] = {0};
        float width = range[axis] / kBuckets;
        for (unsigned i = start; i < end; i++) {
            float q = (p[i].c[axis] - b.mn[axis]) / width;
            unsigned ib = (q != q) ? 0u : static_cast<unsigned>(q); // NaN only if width == 0
            if (ib == kBuckets)
                ib = kBuckets - 1;
            bk_box[ib] = join(bk_box[ib], p[i].bound);
            bk_n[ib]++;
        }
        float cost[kBuckets];
        bucket_costs(bk_box, bk_n, b, cost);
        // The reference starts its minimum search at bucket 1 with bucket 0's cost.
        unsigned best = 1;
        float best_cost = cost[0];
        for (unsigned i = 1; i < kBuckets - 1; i++)
            if (cost[i] < best_cost) {
                best = i;
                best_cost = cost[i];
            }
        if (!(best_cost < cost_leaf || n > kLeafMax)) {
            emit_leaf(b, start, n, depth);
            return;
        }
        unsigned mid;
        if (depth > log2n) {
            std::sort(p.begin() + start, p.begin() + end,
                      [axis](BuildPrim const &x, BuildPrim const &y) { return x.c[axis] < y.c[axis]; });
            mid = (start + end) >> 1;
        } else {
            float split = b.mn[axis] + (best + 1) * range[axis] / kBuckets;
            auto it = std::partition(p.begin() + start, p.begin() + end,
                                     [axis, split](BuildPrim const &x) { return x.c[axis] < split; });
            mid = static_cast<unsigned>(it - p.begin());
            if (mid == start)
                mid++;
            else if (mid == end)
                mid--;
        }
        size_t self = out.size();
        out.push_back(FlatNode{});
        build(start, mid, depth + 1);
        FlatNode &f = out[self];
        std::memcpy(f.mn, b.mn, sizeof f.mn);
        std::memcpy(f.mx, b.mx, sizeof f.mx);
        f.num_prims = 0;
        f.split_axis = axis;
        f.next_child = static_cast<uint32_t>(out.size());
        f.prim_start = 0xFFFFFFFFu;
        build(mid, end, depth + 1);
    }
};

template <class T> std::string device_copy(void **dst, std::vector<T> const &v, uint64_t *bytes) {
    size_t n = std::max<size_t>(v.size() * sizeof(T), 256); // never empty: kernels may read element 0
    if (hipMalloc(dst, n) != hipSuccess)
        return "hipMalloc failed (" + std::to_string(n) + " bytes)";
    if (!v.empty() && hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return "hipMemcpy to device failed";
    *bytes += n;
    return "";
}

struct F4 {
    float x, y, z, w;
};
inline float bits_f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

} // namespace

namespace {

// e8util::mat44 element (row i, column j) of a column-major array.
inline float at(float const m[16], unsigned i, unsigned j) { return m[4 * j + i]; }

// Determinant of the submatrix of m on `rows` x `cols` (ascending index lists) by cofactor
// expansion along its first column, summed from 0 in row order with alternating signs: the
// recursion of mat::det(next_row, curr_col, removed_col) (tensor.h:445-467).
float cofactor_det(float const m[16], unsigned const *rows, unsigned const *cols, unsigned n) {
    if (n == 1)
        return at(m, rows[0], cols[0]);
    float d = 0.0f;
    unsigned sub[4];
    for (unsigned i = 0; i < n; i++) {
        for (unsigned r = 0, k = 0; r < n; r++)
            if (r != i)
                sub[k++] = rows[r];
        float minor = cofactor_det(m, sub, cols + 1, n - 1);
        float a = at(m, rows[i], cols[0]);
        d += (i & 1 ? -a : a) * minor;
    }
    return d;
}

// mat * vec4 (tensor.h:426-435), then vec4::cart (tensor.h:261-268).
HostVec3 apply(float const m[16], HostVec3 p) {
    float h[4] = {p.x, p.y, p.z, 1.0f}, r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (unsigned j = 0; j < 4; j++)
        for (unsigned i = 0; i < 4; i++)
            r[i] = r[i] + at(m, i, j) * h[j];
    float inv_w = 1.0f / r[3];
    return {inv_w * r[0], inv_w * r[1], inv_w * r[2]};
}

} // namespace

std::string transform_mesh(float const m[16], std::vector<HostVec3> &verts, std::vector<HostVec3> &normals) {
    // mat::adjugate (tensor.h:478-497): adj(j, i) = (-1)^(i+j) minor(i, j)
    float adj[16];
    for (unsigned j = 0; j < 4; j++)
        for (unsigned i = 0; i < 4; i++) {
            unsigned rows[3], cols[3];
            for (unsigned r = 0, k = 0; r < 4; r++)
                if (r != i)
                    rows[k++] = r;
            for (unsigned c = 0, k = 0; c < 4; c++)
                if (c != j)
                    cols[k++] = c;
            float mij = cofactor_det(m, rows, cols, 3);
            adj[4 * i + j] = (i + j) & 1 ? -mij : mij; // column i, row j
        }
    // mat::det(adj) (tensor.h:470-476) and operator^(-1) (tensor.h:511-535)
    float d = 0.0f;
    for (unsigned i = 0; i < 4; i++)
        d += at(m, i, 0) * adj[4 * i + 0];
    if (std::fabs(d - 0.0f) <= 1e-5f)
        return "transform is not invertible";
    float k = 1.0f / d, nt[16];
    for (unsigned j = 0; j < 4; j++)
        for (unsigned i = 0; i < 4; i++)
            nt[4 * i + j] = k * adj[4 * j + i]; // transpose of the inverse
    for (HostVec3 &v : verts)
        v = apply(m, v);
    for (HostVec3 &n : normals)
        n = apply(nt, n);
    return "";
}

void Scene::grow_space(std::vector<HostVec3> const &verts) {
    for (HostVec3 const &v : verts) {
        float c[3] = {v.x, v.y, v.z};
        for (int a = 0; a < 3; a++) {
            space_lo[a] = space_empty ? c[a] : std::min(space_lo[a], c[a]);
            space_hi[a] = space_empty ? c[a] : std::max(space_hi[a], c[a]);
        }
        space_empty = false;
    }
}

void Scene::release() {
    void *ptrs[] = {buf.nodes, buf.nodes4, buf.tris, buf.tri_attr, buf.mats, buf.lights, buf.light_cum_power,
                    buf.light_tris, buf.light_tri_cum};
    for (void *q : ptrs)
        if (q)
            (void)hipFree(q);
    buf = DeviceBuffers{};
    dev = DevScene{};
    flat.clear();
    prims.clear();
    committed = false;
}

// The actuators' maps in id order (bvh_path_space_layout::commit, pathspace.cpp:292-309;
// basic_light_sources::commit, lightsources.cpp:45-52): materials, then meshes with their
// material resolved (a missing one is the fail-safe, if_material_container::find,
// materialcontainer.cpp:7-17), then lights on their geometry's mesh.
std::string Scene::flatten_keyed() {
    materials.clear();
    meshes.clear();
    lights.clear();
    geo_ids.clear();
    mat_ids.clear();
    std::map<uint64_t, int32_t> mat_index, geo_index;
    for (auto const &kv : k_mats) {
        mat_index[kv.first] = static_cast<int32_t>(materials.size());
        mat_ids.push_back(kv.first);
        materials.push_back(kv.second);
    }
    for (auto const &kv : k_geos) {
        HostMesh m = kv.second.mesh;
        auto it = kv.second.has_material ? mat_index.find(kv.second.material) : mat_index.end();
        m.material = it == mat_index.end() ? -1 : it->second;
        geo_index[kv.first] = static_cast<int32_t>(meshes.size());
        geo_ids.push_back(kv.first);
        meshes.push_back(std::move(m));
    }
    for (auto const &kv : k_lights) {
        if (kv.second.sky) {
            HostLight l;
            l.mesh = -1;
            for (int c = 0; c < 3; c++)
                l.rad[c] = kv.second.rad[c];
            lights.push_back(l);
            continue;
        }
        auto it = geo_index.find(kv.second.geometry);
        if (it == geo_index.end())
            return "light " + std::to_string(kv.first) + ": its geometry " + std::to_string(kv.second.geometry) +
                   " is not loaded";
        HostLight l;
        l.mesh = it->second;
        for (int c = 0; c < 3; c++)
            l.rad[c] = kv.second.rad[c];
        lights.push_back(l);
    }
    return "";
}

Scene::~Scene() { release(); }

void Scene::build_bvh() {
    std::vector<BuildPrim> p;
    size_t total = 0;
    for (auto const &m : meshes)
        total += m.tris.size() / 3;
    p.reserve(total);
    for (uint32_t g = 0; g < meshes.size(); g++) {
        HostMesh const &m = meshes[g];
        for (uint32_t t = 0; t < m.tris.size() / 3; t++) {
            BuildPrim bp; // primitive_details (pathspace.h:97-113)
            bp.ref = {g, t};
            bp.bound = join_point(bp.bound, m.verts[m.tris[3 * t]]);
            bp.bound = join_point(bp.bound, m.verts[m.tris[3 * t + 1]]);
            bp.bound = join_point(bp.bound, m.verts[m.tris[3 * t + 2]]);
            for (int i = 0; i < 3; i++)
                bp.c[i] = (bp.bound.mx[i] + bp.bound.mn[i]) / 2.0f;
            p.push_back(bp);
        }
    }
    flat.clear();
    flat.reserve(2 * p.size());
    Builder bld{p, flat, std::log2(static_cast<double>(p.size()))};
    if (!p.empty())
        bld.build(0, static_cast<unsigned>(p.size()), 0);
    max_leaf_depth = bld.max_depth;
    prims.resize(p.size());
    for (size_t i = 0; i < p.size(); i++)
        prims[i] = p[i].ref;
}

std::string Scene::commit() {
    if (committed && !keyed)
        return "scene already committed";
    auto t0 = std::chrono::steady_clock::now();
    if (keyed) {
        release();
        std::string e = flatten_keyed();
        if (!e.empty())
            return e;
    }
    size_t total_tris = 0;
    for (auto const &m : meshes)
        total_tris += m.tris.size() / 3;
    if (total_tris >= kMaxPrims)
        return "too many triangles (limit 2^28)";
    for (auto &m : meshes)
        m.light = -1;
    for (size_t l = 0; l < lights.size(); l++) {
        if (lights[l].mesh < 0)
            continue; // the sky has no geometry
        HostMesh &m = meshes[lights[l].mesh];
        if (m.light >= 0)
            return "mesh " + std::to_string(lights[l].mesh) + " carries two lights";
        if (m.tris.empty())
            return "light mesh " + std::to_string(lights[l].mesh) + " has no triangles";
        m.light = static_cast<int32_t>(l);
    }
    // Fail-safe material (mat_fail_safe, material.cpp:54-76): Lambertian 0.8 for meshes
    // without a material.
    DevMaterial fs{};
    fs.kind = kMatFailSafe;
    fs.albedo[0] = fs.albedo[1] = fs.albedo[2] = 0.8f;
    fs.on_a = 1.0f;
    fs.on_b = 0.0f;
    uint32_t fail_safe = static_cast<uint32_t>(materials.size());
    materials.push_back(fs);

    build_bvh();
    auto t1 = std::chrono::steady_clock::now();
    build_seconds = std::chrono::duration<double>(t1 - t0).count();

    std::string err = upload();
    upload_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    if (!err.empty())
        return err;
    dev.fail_safe_mat = fail_safe;
    committed = true;
    return "";
}

std::string Scene::upload() {
    // Interior nodes are numbered by their order in the flattened array.
    std::vector<uint32_t> interior_index(flat.size(), 0xFFFFFFFFu);
    num_interior = 0;
    for (size_t n = 0; n < flat.size(); n++)
        if (flat[n].num_prims == 0)
            interior_index[n] = num_interior++;
    auto ref_of = [&](uint32_t n) -> uint32_t {
        FlatNode const &f = flat[n];
        return f.num_prims > 0 ? leaf_ref(f.prim_start, f.num_prims) : interior_index[n];
    };
    std::vector<F4> nodes(4 * static_cast<size_t>(num_interior));
    for (size_t n = 0; n < flat.size(); n++) {
        if (flat[n].num_prims > 0)
            continue;
        FlatNode const &l = flat[n + 1];
        FlatNode const &r = flat[flat[n].next_child];
        F4 *q = &nodes[4 * static_cast<size_t>(interior_index[n])];
        q[0] = {l.mn[0], l.mx[0], l.mn[1], l.mx[1]};
        q[1] = {r.mn[0], r.mx[0], r.mn[1], r.mx[1]};
        q[2] = {l.mn[2], l.mx[2], r.mn[2], r.mx[2]};
        q[3] = {bits_f(ref_of(static_cast<uint32_t>(n + 1))), bits_f(ref_of(flat[n].next_child)), 0.0f, 0.0f};
    }
    // Wide traversal tree over the reference's leaves. For a ray whose slab test cannot meet NaN
    // the reference reaches exactly the leaves whose own box passes the test: an inner box
    // contains its descendants' boxes, and the entry distance only shrinks and the exit distance
    // only grows with the box (subtraction and multiplication round monotonically), so every
    // ancestor of a passing leaf passes too. Which leaves a ray reaches therefore does not depend
    // on the tree above the leaves, and neither does the closest hit (its tie rule names leaves,
    // not tree positions) or an any-hit answer. The tree above the leaves is rebuilt here for
    // traversal cost: a binned SAH over the leaf boxes on all three axes (the reference bins the
    // widest axis only and stops at four triangles per leaf), inner boxes the exact unions of
    // their leaves' boxes. Then each wide node takes a binary node and repeatedly replaces its
    // interior child of largest surface area by that child's two children, up to kTreeWidth
    // children.
    struct UNode {
        float mn[3], mx[3];
        uint32_t left, right; // children in `up`; leaf nodes: kDone4
        uint32_t ref;         // leaf nodes: the reference leaf's ref
    };
    std::vector<UNode> up;
    {
        struct ULeaf {
            float mn[3], mx[3], c[3];
            uint32_t ref;
            float w; // triangles: the SAH weight
        };
        std::vector<ULeaf> lv;
        for (auto const &f : flat)
            if (f.num_prims > 0) {
                ULeaf l;
                for (int i = 0; i < 3; i++) {
                    l.mn[i] = f.mn[i];
                    l.mx[i] = f.mx[i];
                    l.c[i] = 0.5f * (f.mn[i] + f.mx[i]);
                }
                l.ref = leaf_ref(f.prim_start, f.num_prims);
                l.w = static_cast<float>(f.num_prims);
                lv.push_back(l);
            }
        up.reserve(2 * lv.size());
        constexpr int kBins = 128;
        // bin tables, filled and read before a node's recursion (shared by all levels)
        struct Bins {
            float bmn[kBins][3], bmx[kBins][3];
            double bw[kBins], ra[kBins], rw[kBins];
        };
        std::vector<Bins> binbuf(1);
        Bins &B = binbuf[0];
        auto area = [](const float *mn, const float *mx) {
            double dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
            return dx * dy + dy * dz + dz * dx;
        };
        const uint32_t median_depth = 2 * static_cast<uint32_t>(std::log2(std::max<size_t>(lv.size(), 2))) + 8;
        std::function<uint32_t(uint32_t, uint32_t, uint32_t)> build = [&](uint32_t b, uint32_t e,
                                                                         uint32_t depth) -> uint32_t {
            uint32_t idx = static_cast<uint32_t>(up.size());
            up.push_back(UNode{});
            float mn[3], mx[3], cmn[3], cmx[3];
            for (int i = 0; i < 3; i++) {
                mn[i] = cmn[i] = std::numeric_limits<float>::infinity();
                mx[i] = cmx[i] = -std::numeric_limits<float>::infinity();
            }
            for (uint32_t k = b; k < e; k++)
                for (int i = 0; i < 3; i++) {
                    mn[i] = std::min(mn[i], lv[k].mn[i]);
                    mx[i] = std::max(mx[i], lv[k].mx[i]);
                    cmn[i] = std::min(cmn[i], lv[k].c[i]);
                    cmx[i] = std::max(cmx[i], lv[k].c[i]);
                }
            UNode n;
            std::copy(mn, mn + 3, n.mn);
            std::copy(mx, mx + 3, n.mx);
            n.left = n.right = kDone4;
            n.ref = kDone4;
            if (e - b == 1) {
                n.ref = lv[b].ref;
                std::copy(lv[b].mn, lv[b].mn + 3, n.mn); // the reference leaf's own box
                std::copy(lv[b].mx, lv[b].mx + 3, n.mx);
                up[idx] = n;
                return idx;
            }
            int best_axis = -1, best_bin = 0;
            double best_cost = std::numeric_limits<double>::infinity();
            auto bin_of = [&](ULeaf const &l, int a) {
                float f = (l.c[a] - cmn[a]) / (cmx[a] - cmn[a]) * kBins;
                return !(f >= 0.0f) ? 0 : f >= kBins ? kBins - 1 : static_cast<int>(f); // NaN: bin 0
            };
            if (depth < median_depth)
                for (int a = 0; a < 3; a++) {
                    if (!(cmx[a] > cmn[a]))
                        continue;
                    auto &bmn = B.bmn;
                    auto &bmx = B.bmx;
                    auto &bw = B.bw;
                    auto &ra = B.ra;
                    auto &rw = B.rw;
                    for (int q = 0; q < kBins; q++) {
                        bw[q] = 0.0;
                        for (int i = 0; i < 3; i++) {
                            bmn[q][i] = std::numeric_limits<float>::infinity();
                            bmx[q][i] = -std::numeric_limits<float>::infinity();
                        }
                    }
                    for (uint32_t k = b; k < e; k++) {
                        int q = bin_of(lv[k], a);
                        bw[q] += lv[k].w;
                        for (int i = 0; i < 3; i++) {
                            bmn[q][i] = std::min(bmn[q][i], lv[k].mn[i]);
                            bmx[q][i] = std::max(bmx[q][i], lv[k].mx[i]);
                        }
                    }
                    // suffix sweep: right-side area and weight of a split after bin q
                    float smn[3], smx[3];
                    for (int i = 0; i < 3; i++) {
                        smn[i] = std::numeric_limits<float>::infinity();
                        smx[i] = -std::numeric_limits<float>::infinity();
                    }
                    double sw = 0;
                    for (int q = kBins - 1; q > 0; q--) {
                        for (int i = 0; i < 3; i++) {
                            smn[i] = std::min(smn[i], bmn[q][i]);
                            smx[i] = std::max(smx[i], bmx[q][i]);
                        }
                        sw += bw[q];
                        ra[q - 1] = sw > 0 ? area(smn, smx) : 0.0;
                        rw[q - 1] = sw;
                    }
                    for (int i = 0; i < 3; i++) {
                        smn[i] = std::numeric_limits<float>::infinity();
                        smx[i] = -std::numeric_limits<float>::infinity();
                    }
                    sw = 0;
                    for (int q = 0; q < kBins - 1; q++) {
                        for (int i = 0; i < 3; i++) {
                            smn[i] = std::min(smn[i], bmn[q][i]);
                            smx[i] = std::max(smx[i], bmx[q][i]);
                        }
                        sw += bw[q];
                        if (sw <= 0 || rw[q] <= 0)
                            continue;
                        double c = area(smn, smx) * sw + ra[q] * rw[q];
                        if (c < best_cost) {
                            best_cost = c;
                            best_axis = a;
                            best_bin = q;
                        }
                    }
                }
            uint32_t mid;
            if (best_axis >= 0) {
                mid = static_cast<uint32_t>(
                    std::partition(lv.begin() + b, lv.begin() + e,
                                   [&](ULeaf const &l) { return bin_of(l, best_axis) <= best_bin; }) -
                    lv.begin());
            } else {
                // all centroids equal, or too deep: median of the widest centroid axis
                int a = 0;
                for (int i = 1; i < 3; i++)
                    if (cmx[i] - cmn[i] > cmx[a] - cmn[a])
                        a = i;
                mid = (b + e) / 2;
                std::nth_element(lv.begin() + b, lv.begin() + mid, lv.begin() + e,
                                 [a](ULeaf const &x, ULeaf const &y) { return x.c[a] < y.c[a]; });
            }
            if (mid == b || mid == e)
                mid = (b + e) / 2;
            uint32_t l = build(b, mid, depth + 1);
            uint32_t r = build(mid, e, depth + 1);
            n.left = l;
            n.right = r;
            up[idx] = n;
            return idx;
        };
        if (!lv.empty())
            build(0, static_cast<uint32_t>(lv.size()), 0);
    }
    // Record (kWideRows x float4): for each plane (min x, max x, min y, max y, min z, max z) the
    // children in groups of four, then the child refs in groups of four (interior: wide node
    // index; kDone4: no child), padding.
    std::vector<F4> nodes4;
    uint32_t max_push4 = 0;
    std::function<uint32_t(uint32_t, uint32_t)> collapse = [&](uint32_t n, uint32_t pushes) -> uint32_t {
        UNode const &f = up[n];
        if (f.ref != kDone4) {
            max_push4 = std::max(max_push4, pushes);
            return f.ref;
        }
        uint32_t c[kTreeWidth] = {f.left, f.right};
        uint32_t k = 2;
        while (k < kTreeWidth) {
            int best = -1;
            float best_area = -1.0f;
            for (uint32_t i = 0; i < k; i++) {
                UNode const &q = up[c[i]];
                if (q.ref != kDone4)
                    continue;
                float dx = q.mx[0] - q.mn[0], dy = q.mx[1] - q.mn[1], dz = q.mx[2] - q.mn[2];
                float a = dx * dy + dy * dz + dz * dx;
                if (a > best_area) {
                    best_area = a;
                    best = static_cast<int>(i);
                }
            }
            if (best < 0)
                break;
            uint32_t m = c[best];
            c[best] = up[m].left;
            c[k++] = up[m].right;
        }
        // Children stored smallest surface area first. Any-hit traversal descends into the first
        // hit child, and the closest-hit sort keeps stored order among equal entry distances
        // (rays starting inside several boxes all enter them at tmin), so the smaller boxes,
        // whose geometry lies nearest such a ray, are tried first (measured: shadow rays visit
        // 10% fewer nodes, extension rays 2.5% fewer).
        auto sa = [&](uint32_t i) {
            UNode const &q = up[i];
            float dx = q.mx[0] - q.mn[0], dy = q.mx[1] - q.mn[1], dz = q.mx[2] - q.mn[2];
            return dx * dy + dy * dz + dz * dx;
        };
        std::stable_sort(c, c + k, [&](uint32_t x, uint32_t y) { return sa(x) < sa(y); });
        uint32_t idx = static_cast<uint32_t>(nodes4.size() / kWideRows);
        nodes4.resize(nodes4.size() + kWideRows);
        uint32_t refs[kTreeWidth];
        for (uint32_t i = 0; i < kTreeWidth; i++)
            refs[i] = kDone4;
        for (uint32_t i = 0; i < k; i++)
            refs[i] = collapse(c[i], pushes + k - 1);
        F4 *q = &nodes4[kWideRows * static_cast<size_t>(idx)];
        for (uint32_t g = 0; g < kWideGroups; g++) {
            for (int a = 0; a < 3; a++) {
                float lo[4], hi[4];
                for (uint32_t j = 0; j < 4; j++) {
                    uint32_t i = 4 * g + j;
                    // an empty slot gets the box [+inf, +inf]: every wide-tree ray misses it
                    lo[j] = i < k ? up[c[i]].mn[a] : std::numeric_limits<float>::infinity();
                    hi[j] = i < k ? up[c[i]].mx[a] : std::numeric_limits<float>::infinity();
                }
                q[(2 * a) * kWideGroups + g] = {lo[0], lo[1], lo[2], lo[3]};
                q[(2 * a + 1) * kWideGroups + g] = {hi[0], hi[1], hi[2], hi[3]};
            }
            q[6 * kWideGroups + g] = {bits_f(refs[4 * g]), bits_f(refs[4 * g + 1]), bits_f(refs[4 * g + 2]),
                                      bits_f(refs[4 * g + 3])};
        }
        return idx;
    };
    uint32_t root4 = up.empty() ? kEmptyRef : collapse(0, 0);
    num_nodes4 = static_cast<uint32_t>(nodes4.size() / kWideRows);
    max_stack4 = max_push4 + 1;
    if (kTopNodes > 0 && root4 != kEmptyRef && !(root4 & kLeafBit)) {
        // The first kTopNodes nodes in breadth-first order (the top of the tree, which traversal
        // kernels may keep in LDS) take indices 0.., the rest keep their depth-first order.
        std::vector<uint32_t> order, renum(num_nodes4, kDone4);
        order.push_back(root4);
        renum[root4] = 0;
        for (size_t h = 0; h < order.size() && order.size() < kTopNodes; h++)
            for (uint32_t g = 0; g < kWideGroups; g++) {
                F4 const &q = nodes4[kWideRows * static_cast<size_t>(order[h]) + 6 * kWideGroups + g];
                const float fr[4] = {q.x, q.y, q.z, q.w};
                for (float f : fr) {
                    uint32_t c;
                    std::memcpy(&c, &f, 4);
                    if (c != kDone4 && !(c & kLeafBit) && order.size() < kTopNodes && renum[c] == kDone4) {
                        renum[c] = static_cast<uint32_t>(order.size());
                        order.push_back(c);
                    }
                }
            }
        uint32_t next = static_cast<uint32_t>(order.size());
        for (uint32_t o = 0; o < num_nodes4; o++)
            if (renum[o] == kDone4)
                renum[o] = next++;
        std::vector<F4> moved(nodes4.size());
        for (uint32_t o = 0; o < num_nodes4; o++) {
            F4 *dst = &moved[kWideRows * static_cast<size_t>(renum[o])];
            std::copy(&nodes4[kWideRows * static_cast<size_t>(o)], &nodes4[kWideRows * static_cast<size_t>(o)] + kWideRows,
                      dst);
            for (uint32_t g = 0; g < kWideGroups; g++) {
                F4 &q = dst[6 * kWideGroups + g];
                float *fr[4] = {&q.x, &q.y, &q.z, &q.w};
                for (float *f : fr) {
                    uint32_t c;
                    std::memcpy(&c, f, 4);
                    if (c != kDone4 && !(c & kLeafBit))
                        *f = bits_f(renum[c]);
                }
            }
        }
        nodes4.swap(moved);
        root4 = renum[root4];
    }
    {
        // Consistency of the wide tree before anything traverses it: refs in range, every
        // binary leaf reached exactly once, every node reached from the root.
        std::vector<uint8_t> seen_leaf(prims.size(), 0), seen_node(num_nodes4, 0);
        size_t leaves = 0, binary_leaves = 0;
        for (auto const &f : flat)
            binary_leaves += f.num_prims > 0 ? 1 : 0;
        std::vector<uint32_t> todo;
        if (root4 != kEmptyRef)
            todo.push_back(root4);
        while (!todo.empty()) {
            uint32_t r = todo.back();
            todo.pop_back();
            if (r & kLeafBit) {
                uint32_t start = r & 0x0FFFFFFFu, cnt = ((r >> 28) & 3u) + 1u;
                if (start + cnt > prims.size() || seen_leaf[start])
                    return "internal error: wide tree leaf out of range or repeated";
                seen_leaf[start] = 1;
                leaves++;
                continue;
            }
            if (r >= num_nodes4 || seen_node[r])
                return "internal error: wide tree node out of range or repeated";
            seen_node[r] = 1;
            for (uint32_t g = 0; g < kWideGroups; g++) {
                F4 const &q = nodes4[kWideRows * static_cast<size_t>(r) + 6 * kWideGroups + g];
                const float fr[4] = {q.x, q.y, q.z, q.w};
                for (float f : fr) {
                    uint32_t c;
                    std::memcpy(&c, &f, 4);
                    if (c != kDone4)
                        todo.push_back(c);
                }
            }
        }
        if (leaves != binary_leaves || std::count(seen_node.begin(), seen_node.end(), 1) != num_nodes4)
            return "internal error: wide tree does not cover the binary tree";
    }

    std::vector<F4> tris(3 * prims.size()), attr(3 * prims.size());
    for (size_t i = 0; i < prims.size(); i++) {
        HostMesh const &m = meshes[prims[i].mesh];
        uint32_t t = prims[i].tri;
        HostVec3 a = m.verts[m.tris[3 * t]], b = m.verts[m.tris[3 * t + 1]], c = m.verts[m.tris[3 * t + 2]];
        HostVec3 na = m.normals[m.tris[3 * t]], nb = m.normals[m.tris[3 * t + 1]],
                 nc = m.normals[m.tris[3 * t + 2]];
        uint32_t mat = m.material >= 0 ? static_cast<uint32_t>(m.material) : static_cast<uint32_t>(materials.size() - 1);
        tris[3 * i] = {a.x, a.y, a.z, bits_f(prims[i].mesh)};
        tris[3 * i + 1] = {b.x, b.y, b.z, bits_f(t)};
        tris[3 * i + 2] = {c.x, c.y, c.z, 0.0f};
        attr[3 * i] = {na.x, na.y, na.z, bits_f(mat)};
        attr[3 * i + 1] = {nb.x, nb.y, nb.z, bits_f(static_cast<uint32_t>(m.light))};
        attr[3 * i + 2] = {nc.x, nc.y, nc.z, 0.0f};
    }
    // Lights: per-mesh area CDF (geometry.cpp:105-118) and power CDF (lightsources.cpp:45-52),
    // power = pi * area * radiance (light.cpp:20-23).
    std::vector<DevLight> dl;
    std::vector<float> cum_power, tri_cum;
    std::vector<F4> ltris;
    float total_power = 0;
    for (auto const &L : lights) {
        if (L.mesh < 0) {
            // sky_light::set_scene_boundary (light.cpp:96-100) applied at commit with the path
            // space's bounds: m_dia = 2 * enclosing_radius = 2 * (|max - min| / 2), m_ref_p =
            // (max + min) / 2 with z = 0 (tensor.h:811-813); power = pi * dia^2 / 2 per channel
            // (light.cpp:151).
            DevLight d{};
            d.kind = kLightSky;
            d.mesh = kNoMesh;
            float dx = space_hi[0] - space_lo[0], dy = space_hi[1] - space_lo[1], dz = space_hi[2] - space_lo[2];
            float nn = 0.0f;
            nn += dx * dx;
            nn += dy * dy;
            nn += dz * dz;
            d.sky_dia = 2 * (std::sqrt(nn) / 2);
            d.sky_ref[0] = (space_hi[0] + space_lo[0]) / 2;
            d.sky_ref[1] = (space_hi[1] + space_lo[1]) / 2;
            d.sky_ref[2] = 0.0f;
            float p = static_cast<float>(M_PI) * (d.sky_dia * d.sky_dia / 2);
            float pn = 0.0f;
            pn += p * p;
            pn += p * p;
            pn += p * p;
            total_power += std::sqrt(pn);
            cum_power.push_back(total_power);
            for (int c = 0; c < 3; c++)
                d.rad[c] = L.rad[c];
            d.prob = std::sqrt(pn); // divided by the total below
            dl.push_back(d);
            continue;
        }
        HostMesh const &m = meshes[L.mesh];
        DevLight d{};
        d.tri_begin = static_cast<uint32_t>(tri_cum.size());
        d.tri_count = static_cast<uint32_t>(m.tris.size() / 3);
        float cum = 0;
        for (uint32_t t = 0; t < d.tri_count; t++) {
            HostVec3 a = m.verts[m.tris[3 * t]], b = m.verts[m.tris[3 * t + 1]], c = m.verts[m.tris[3 * t + 2]];
            HostVec3 na = m.normals[m.tris[3 * t]], nb = m.normals[m.tris[3 * t + 1]],
                     nc = m.normals[m.tris[3 * t + 2]];
            float ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
            float vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
            float cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
            float nn = 0.0f;
            nn += cx * cx;
            nn += cy * cy;
            nn += cz * cz;
            float area = 0.5f * std::sqrt(nn);
            cum += area;
            tri_cum.push_back(cum);
            ltris.push_back({a.x, a.y, a.z, 0});
            ltris.push_back({b.x, b.y, b.z, 0});
            ltris.push_back({c.x, c.y, c.z, 0});
            ltris.push_back({na.x, na.y, na.z, 0});
            ltris.push_back({nb.x, nb.y, nb.z, 0});
            ltris.push_back({nc.x, nc.y, nc.z, 0});
        }
        d.area = cum;
        float k = static_cast<float>(M_PI) * cum;
        float px = k * L.rad[0], py = k * L.rad[1], pz = k * L.rad[2];
        float pn = 0.0f;
        pn += px * px;
        pn += py * py;
        pn += pz * pz;
        total_power += std::sqrt(pn);
        cum_power.push_back(total_power);
        d.rad[0] = L.rad[0];
        d.rad[1] = L.rad[1];
        d.rad[2] = L.rad[2];
        d.mesh = static_cast<uint32_t>(L.mesh);
        d.kind = kLightArea;
        d.prob = std::sqrt(pn); // divided by the total below
        dl.push_back(d);
    }
    // basic_light_sources::sample_light's probability (lightsources.cpp:73); lights of zero
    // total power get probability 0 (the reference's 0 / 0 would make every estimate NaN)
    for (auto &d : dl)
        d.prob = total_power > 0.0f ? d.prob / total_power : 0.0f;

    std::string e;
    buf.bytes = 0;
    if (!(e = device_copy(&buf.nodes, nodes, &buf.bytes)).empty() ||
        !(e = device_copy(&buf.nodes4, nodes4, &buf.bytes)).empty() ||
        !(e = device_copy(&buf.tris, tris, &buf.bytes)).empty() ||
        !(e = device_copy(&buf.tri_attr, attr, &buf.bytes)).empty() ||
        !(e = device_copy(&buf.mats, materials, &buf.bytes)).empty() ||
        !(e = device_copy(&buf.lights, dl, &buf.bytes)).empty() ||
        !(e = device_copy(&buf.light_cum_power, cum_power, &buf.bytes)).empty() ||
        !(e = device_copy(&buf.light_tris, ltris, &buf.bytes)).empty() ||
        !(e = device_copy(&buf.light_tri_cum, tri_cum, &buf.bytes)).empty())
        return e;

    dev.nodes = buf.nodes;
    dev.nodes4 = buf.nodes4;
    dev.root4 = root4;
    dev.num_nodes4 = num_nodes4;
    dev.tris = buf.tris;
    dev.tri_attr = buf.tri_attr;
    dev.mats = static_cast<const DevMaterial *>(buf.mats);
    dev.lights = static_cast<const DevLight *>(buf.lights);
    dev.light_cum_power = static_cast<const float *>(buf.light_cum_power);
    dev.light_tris = buf.light_tris;
    dev.light_tri_cum = static_cast<const float *>(buf.light_tri_cum);
    dev.root_ref = flat.empty() ? kEmptyRef : ref_of(0);
    dev.num_lights = static_cast<uint32_t>(lights.size());
    dev.num_mats = static_cast<uint32_t>(materials.size());
    dev.total_power = total_power;
    dev.max_stack = std::max(max_leaf_depth + 1, max_stack4);
    dev.num_prims = static_cast<uint32_t>(prims.size());
    dev.num_interior = num_interior;
    // the wide tree's rows are addressed by 32-bit byte offsets: a larger tree is not used
    dev.far_coords = static_cast<uint64_t>(num_nodes4) * kWideRows * sizeof(F4) < (1ull << 32) ? 0u : 1u;
    for (int a = 0; a < 3; a++) {
        dev.space_lo[a] = space_lo[a];
        dev.space_range[a] = space_hi[a] - space_lo[a];
    }
    for (int a = 0; a < 3; a++) {
        float lo = flat.empty() ? 0.0f : flat[0].mn[a], hi = flat.empty() ? 0.0f : flat[0].mx[a];
        dev.box_lo[a] = lo;
        dev.box_inv[a] = hi > lo ? 1.0f / (hi - lo) : 0.0f;
        // every node box lies inside the root box
        if (!(std::fabs(lo) <= 1e30f && std::fabs(hi) <= 1e30f))
            dev.far_coords = 1;
    }
    return "";
}

} // namespace x3d
