// BVH traversal on gfx950: closest hit and any hit over the reference layout.
//
// One lane per ray. Each interior visit reads one 64-byte node holding both child boxes, tests
// both with the reference's slab test, descends into the nearer hit child and pushes the other
// (with its entry distance) on a per-lane stack. The stack's first LDS_DEPTH entries live in
// LDS, entry-major ([entry][lane]) so that lanes of a wave hit distinct banks; deeper entries
// spill to a per-thread slice of global memory sized from the tree depth.
//
// Loop structure (while-while with postponed leaves, Aila & Laine, HPG 2009): a lane that reaches
// a leaf parks it and keeps descending until every active lane of the wave holds a leaf; then
// all lanes test their leaves' triangles together, so the expensive triangle tests run with the
// wave's lanes converged.
//
// Closest hit returns exactly the reference's winner (bvh_path_space_layout::intersect,
// pathspace.cpp:330-413): the reference visits leaves in decreasing flattened order and keeps
// the first of equal-distance hits, i.e. the hit in the leaf with the largest prim_start, then
// the lowest primitive inside that leaf. That rule is applied explicitly, so the visiting order
// and the distance cull (boxes starting beyond best * (1 + 2^-12) are skipped; the relative
// slack keeps a hit whose FP32 distance falls slightly below its box's entry distance) do not
// change the result. The reference order (right child first, no culling) is available for
// checking.
#pragma once

#include "device_math.h"

namespace x3d {

struct ClosestHit {
    int32_t prim;
    float t, b1, b2;
};

constexpr uint32_t kDone = 0xFFFFFFFFu; // no node left (same bits as kEmptyRef)

template <int WG, int LDS_DEPTH> struct TravStack {
    uint2 *lds;   // this lane's column of the workgroup's [LDS_DEPTH][WG] array
    uint2 *spill; // this thread's global overflow slice
    int sp;
    X3D_DI void push(uint32_t ref, float t) {
        uint2 e = make_uint2(ref, __float_as_uint(t));
        if (sp < LDS_DEPTH)
            lds[sp * WG] = e;
        else
            spill[sp - LDS_DEPTH] = e;
        sp++;
    }
    X3D_DI uint2 pop() {
        sp--;
        if (sp < LDS_DEPTH)
            return lds[sp * WG];
        // A distinct (non-temporal) load keeps the compiler from merging the two reads into one
        // generic-address load through a selected pointer, which would turn every pop into a
        // flat load that waits on both the LDS and the vector memory counters.
        unsigned long long e = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long *>(spill) +
                                                          (sp - LDS_DEPTH));
        return make_uint2(static_cast<uint32_t>(e), static_cast<uint32_t>(e >> 32));
    }
    // Next stacked node whose entry distance does not exceed `cut`, or kDone.
    X3D_DI uint32_t pop_next(float cut) {
        while (sp > 0) {
            uint2 e = pop();
            if (!(__uint_as_float(e.y) > cut))
                return e.x;
        }
        return kDone;
    }
};

X3D_DI bool is_interior(uint32_t ref) { return ref < kLeafBit; }
X3D_DI bool is_leaf_ref(uint32_t ref) { return ref >= kLeafBit && ref != kDone; }
X3D_DI uint32_t leaf_start(uint32_t ref) { return ref & 0x0FFFFFFFu; }
X3D_DI uint32_t leaf_count(uint32_t ref) { return ((ref >> 28) & 3u) + 1u; }

X3D_DI void load_node(DevScene const &S, uint32_t idx, float4 &f0, float4 &f1, float4 &f2, float4 &f3) {
    const float4 *n = static_cast<const float4 *>(S.nodes) + 4u * idx;
    f0 = ldg(n);
    f1 = ldg(n + 1);
    f2 = ldg(n + 2);
    f3 = ldg(n + 3);
}

// Visits interior node `ref`: tests both child boxes against [tmin, tmax], drops children
// entering beyond `cut`, pushes the farther of two hit children and returns the one to descend
// into (or the next stacked node, or kDone). REF_ORDER pushes left, descends right
// (pathspace.cpp:371-381 with its LIFO stack).
template <bool REF_ORDER, int WG, int LDS_DEPTH>
X3D_DI uint32_t visit_interior(DevScene const &S, uint32_t ref, f3 o, f3 vinv, float tmin, float tmax, float cut,
                               TravStack<WG, LDS_DEPTH> &st, bool may_nan) {
    float4 f0, f1, f2, f3n;
    load_node(S, ref, f0, f1, f2, f3n);
    float tl = 0.0f, tr = 0.0f;
    bool hl, hr;
    if (!may_nan) {
        hl = slab_fast(f0.x, f0.y, f0.z, f0.w, f2.x, f2.y, o, vinv, tmin, tmax, &tl);
        hr = slab_fast(f1.x, f1.y, f1.z, f1.w, f2.z, f2.w, o, vinv, tmin, tmax, &tr);
    } else {
        hl = slab(f0.x, f0.y, f0.z, f0.w, f2.x, f2.y, o, vinv, tmin, tmax, &tl);
        hr = slab(f1.x, f1.y, f1.z, f1.w, f2.z, f2.w, o, vinv, tmin, tmax, &tr);
    }
    uint32_t rl = __float_as_uint(f3n.x), rr = __float_as_uint(f3n.y);
    if (!REF_ORDER) {
        hl = hl && !(tl > cut);
        hr = hr && !(tr > cut);
    }
    if (hl && hr) {
        if (REF_ORDER || tr < tl) {
            st.push(rl, tl);
            return rr;
        }
        st.push(rr, tr);
        return rl;
    }
    if (hl)
        return rl;
    if (hr)
        return rr;
    return st.pop_next(REF_ORDER ? __builtin_inff() : cut);
}

// Compare-exchange of two (entry distance, ref) pairs held in named variables, ascending distance.
X3D_DI void cswap2(float &ta, uint32_t &ra, float &tb, uint32_t &rb) {
    bool sw = tb < ta;
    float lo = sw ? tb : ta, hi = sw ? ta : tb;
    uint32_t rl = sw ? rb : ra, rh = sw ? ra : rb;
    ta = lo;
    tb = hi;
    ra = rl;
    rb = rh;
}

// Compare-exchange of two (entry distance, ref) slots, ascending distance.
X3D_DI void cswap(float *t, uint32_t *r, int a, int b) {
    bool sw = t[b] < t[a];
    float ta = sw ? t[b] : t[a], tb = sw ? t[a] : t[b];
    uint32_t ra = sw ? r[b] : r[a], rb = sw ? r[a] : r[b];
    t[a] = ta;
    t[b] = tb;
    r[a] = ra;
    r[b] = rb;
}

// Rows of a wide node that hold a ray's near and far plane of each axis. For a ray whose slab
// test cannot meet NaN (finite, nonzero inverse direction; every box has min <= max, and an
// empty slot's planes are equal), min((lo - o) * vinv, (hi - o) * vinv) is the lo plane's
// value when vinv > 0 and the hi plane's when vinv < 0, and max is the other one (rounding is
// monotone), so loading the planes in that order per ray gives the reference's slab values
// without the per-axis min and max.
struct WideSel {
    uint32_t nx, ny, nz; // byte offsets of the near-plane rows within a wide node; the far-plane
                         // row of an axis is the other row of its pair (offset ^ row bytes)
};

X3D_DI WideSel wide_sel(f3 vinv) {
    constexpr uint32_t g = kWideGroups * 16u; // bytes per row
    WideSel s;
    s.nx = vinv.x < 0.0f ? g : 0u;
    s.ny = vinv.y < 0.0f ? 3u * g : 2u * g;
    s.nz = vinv.z < 0.0f ? 5u * g : 4u * g;
    return s;
}

// Visits wide node `ref` (a ray that cannot meet NaN in the slab test): tests the child boxes,
// drops children entering beyond `cut` (the caller's min(tmax, cull distance): finite), pushes
// the hit children but the nearest, farthest first, and returns the nearest (or the next
// stacked node, or kDone).
// Children with equal entry distances may come in either order; the order never changes a
// result. Unordered (any-hit queries, which stop at the first hit anywhere on the segment):
// no sort; the first hit child in stored order (smallest surface area first) is returned and
// the others are pushed so that they pop in that order.
// Hit and cull in one comparison: without NaN, !(lo > min(hi, tmax)) && !(lo > cull) is
// !(lo > min(hi, min(tmax, cull))).
// `top` (LDS) holds the first kTopNodes nodes when a kernel staged them, else null.
// NEAR_ROWS false: the rows are loaded in stored order and each axis takes the min and max of
// its two plane distances (the same values; for kernels that cannot keep `sel` in registers).
template <bool ORDERED = true, bool NEAR_ROWS = true, int WG, int LDS_DEPTH>
X3D_DI uint32_t visit_wide(DevScene const &S, uint32_t ref, f3 o, f3 vinv, WideSel const &sel, float tmin, float cut,
                           TravStack<WG, LDS_DEPTH> &st, const float4 *top = nullptr) {
    // the node's rows addressed by 32-bit byte offsets from the array base (a wave-uniform base
    // then needs no 64-bit address arithmetic per row; a scene whose wide tree spans 4 GiB or
    // more traverses the binary tree only, DevScene::far_coords)
    const bool in_top = kTopNodes > 0 && top && ref < kTopNodes;
    const char *base = in_top ? reinterpret_cast<const char *>(top) : static_cast<const char *>(S.nodes4);
    const uint32_t nb = ref * (kWideRows * 16u);
    constexpr uint32_t kRowPair = kWideGroups * 16u; // near ^ kRowPair = far row of the same axis
    auto row = [&](uint32_t off) {
        const float4 *q = reinterpret_cast<const float4 *>(base + (nb + off));
        return in_top ? *q : ldg(q);
    };
    // t: entry distance of a hit child, +inf for a miss (a hit child enters at most at
    // min(tmax, cut), which is finite), so a child is valid exactly when t is finite; r: the
    // child refs as stored.
    float t[kTreeWidth];
    uint32_t r[kTreeWidth];
    bool hit[kTreeWidth];
#pragma unroll
    for (uint32_t g = 0; g < kWideGroups; g++) {
        const uint32_t gb = 16u * g;
        const uint32_t ox = NEAR_ROWS ? sel.nx : 0u, oy = NEAR_ROWS ? sel.ny : 2u * kRowPair;
        const uint32_t oz = NEAR_ROWS ? sel.nz : 4u * kRowPair;
        const float4 qnx = row(ox + gb), qfx = row((ox ^ kRowPair) + gb);
        const float4 qny = row(oy + gb), qfy = row((oy ^ kRowPair) + gb);
        const float4 qnz = row(oz + gb), qfz = row((oz ^ kRowPair) + gb);
        const float4 qrf = row(6u * kRowPair + gb);
        const float nx[4] = {qnx.x, qnx.y, qnx.z, qnx.w}, fx[4] = {qfx.x, qfx.y, qfx.z, qfx.w};
        const float ny[4] = {qny.x, qny.y, qny.z, qny.w}, fy[4] = {qfy.x, qfy.y, qfy.z, qfy.w};
        const float nz[4] = {qnz.x, qnz.y, qnz.z, qnz.w}, fz[4] = {qfz.x, qfz.y, qfz.z, qfz.w};
        const float rf[4] = {qrf.x, qrf.y, qrf.z, qrf.w};
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            float ax = (nx[j] - o.x) * vinv.x, bx = (fx[j] - o.x) * vinv.x;
            float ay = (ny[j] - o.y) * vinv.y, by = (fy[j] - o.y) * vinv.y;
            float az = (nz[j] - o.z) * vinv.z, bz = (fz[j] - o.z) * vinv.z;
            if (!NEAR_ROWS) {
                const float mx = fminf(ax, bx), my = fminf(ay, by), mz = fminf(az, bz);
                bx = fmaxf(ax, bx);
                by = fmaxf(ay, by);
                bz = fmaxf(az, bz);
                ax = mx;
                ay = my;
                az = mz;
            }
            const float lo = fmaxf(fmaxf(ax, ay), fmaxf(az, tmin));
            const float hi = fminf(fminf(bx, by), fminf(bz, cut));
            // an empty slot's box [+inf, +inf] always misses (the ray's vinv is finite, nonzero)
            hit[4 * g + j] = !(lo > hi);
            t[4 * g + j] = ORDERED && !hit[4 * g + j] ? __builtin_inff() : lo; // misses sort last
            r[4 * g + j] = __float_as_uint(rf[j]);
        }
    }
    if (!ORDERED) {
        // children in stored order: the first hit one is descended into, the others are pushed
        // last first, so they come off the stack in stored order as well
        uint32_t next = kDone;
        float tn = 0.0f;
#pragma unroll
        for (int i = kTreeWidth - 1; i >= 0; i--)
            if (hit[i]) {
                if (next != kDone)
                    st.push(next, tn);
                next = r[i];
                tn = t[i];
            }
        return next != kDone ? next : st.pop_next(cut);
    }
    // sorting network, ascending entry distance
    if (kTreeWidth == 4) {
        // on named variables: the same network over the arrays let the compiler track the
        // permutation as select chains of indices (about twice the instructions)
        float t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
        uint32_t r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
        cswap2(t0, r0, t1, r1);
        cswap2(t2, r2, t3, r3);
        cswap2(t0, r0, t2, r2);
        cswap2(t1, r1, t3, r3);
        cswap2(t1, r1, t2, r2);
        if (t3 != __builtin_inff())
            st.push(r3, t3);
        if (t2 != __builtin_inff())
            st.push(r2, t2);
        if (t1 != __builtin_inff())
            st.push(r1, t1);
        if (t0 != __builtin_inff())
            return r0;
        return st.pop_next(cut);
    } else {
        constexpr int kNet[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 1}, {2, 3},
                                     {4, 5}, {6, 7}, {2, 4}, {3, 5}, {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
        for (int c = 0; c < 19; c++)
            cswap(t, r, kNet[c][0], kNet[c][1]);
    }
#pragma unroll
    for (int i = kTreeWidth - 1; i > 0; i--)
        if (t[i] != __builtin_inff())
            st.push(r[i], t[i]);
    if (t[0] != __builtin_inff())
        return r[0];
    return st.pop_next(cut);
}

// REF_ORDER: the reference's traversal order without culling.
template <int WG, int LDS_DEPTH, bool REF_ORDER, bool NEAR_ROWS = true>
X3D_DI ClosestHit closest_hit(DevScene const &S, f3 o, f3 v, TravStack<WG, LDS_DEPTH> &st,
                              uint32_t &pops, uint32_t &tests) {
    ClosestHit h{-1, __builtin_inff(), 0.0f, 0.0f};
    if (S.root_ref == kEmptyRef)
        return h;
    f3 vinv = mk(1.0f / v.x, 1.0f / v.y, 1.0f / v.z);
    // the wide tree, unless the ray may meet NaN or the reference order is asked for
    const bool wide = !REF_ORDER && !slab_may_nan(o, vinv, S.far_coords != 0);
    const WideSel sel = NEAR_ROWS ? wide_sel(vinv) : WideSel{};
    const float4 *tris = static_cast<const float4 *>(S.tris);
    uint32_t best_leaf = 0;
    uint32_t ref = wide ? S.root4 : S.root_ref, leaf = kDone;
    if (!is_interior(ref)) {
        leaf = ref;
        ref = kDone;
    }
    // cull distance min(tmax, best t (1 + slack)), kept in step with h.t: children that pass the
    // slab test enter at most at tmax, so the tmax bound changes no cull decision
    float cut = REF_ORDER ? __builtin_inff() : kTMax;
    st.sp = 0;
    for (;;) {
        // Descend; park the first leaf met and keep going until all active lanes hold one.
        while (is_interior(ref)) {
            if (wide) {
                pops += kTreeWidth / 2; // a wide node counts as W/2 binary nodes (same bytes and tests)
                ref = visit_wide<true, NEAR_ROWS>(S, ref, o, vinv, sel, kTMin, cut, st);
            } else {
                pops++;
                ref = visit_interior<REF_ORDER>(S, ref, o, vinv, kTMin, kTMax, cut, st, true);
            }
            if (is_leaf_ref(ref) && leaf == kDone) {
                leaf = ref;
                ref = st.pop_next(cut);
            }
            if (__all(leaf != kDone))
                break;
        }
        // Test the parked leaf, then any leaves that follow on the stack.
        while (leaf != kDone) {
            uint32_t s0 = leaf_start(leaf), cnt = leaf_count(leaf);
            for (uint32_t i = s0; i < s0 + cnt; i++) {
                const float4 *tp = tris + 3 * static_cast<size_t>(i);
                float4 a = ldg(tp), b = ldg(tp + 1), c = ldg(tp + 2);
                float c1, c2, t0;
                tests++;
                if (tri_hit(o, v, xyz(a), xyz(b), xyz(c), kTMin, kTMax, &c1, &c2, &t0) &&
                    (t0 < h.t || (t0 == h.t && s0 > best_leaf))) {
                    h.t = t0;
                    h.b1 = c1;
                    h.b2 = c2;
                    h.prim = static_cast<int32_t>(i);
                    best_leaf = s0;
                    if (!REF_ORDER)
                        cut = fminf(kTMax, t0 + t0 * kCullSlack);
                }
            }
            if (is_leaf_ref(ref)) {
                leaf = ref;
                ref = st.pop_next(cut);
            } else {
                leaf = kDone;
            }
        }
        if (ref == kDone)
            break;
    }
    return h;
}

// has_intersect (pathspace.cpp:415-456): any triangle hit in [tmin, tmax] under boxes clipped to
// the same interval; the result does not depend on the visiting order.
template <int WG, int LDS_DEPTH, bool NEAR_ROWS = true>
X3D_DI bool any_hit(DevScene const &S, f3 o, f3 v, float tmin, float tmax, TravStack<WG, LDS_DEPTH> &st,
                    uint32_t &pops, uint32_t &tests) {
    if (S.root_ref == kEmptyRef)
        return false;
    f3 vinv = mk(1.0f / v.x, 1.0f / v.y, 1.0f / v.z);
    const bool wide = !slab_may_nan(o, vinv, S.far_coords != 0);
    const WideSel sel = NEAR_ROWS ? wide_sel(vinv) : WideSel{};
    const float4 *tris = static_cast<const float4 *>(S.tris);
    uint32_t ref = wide ? S.root4 : S.root_ref, leaf = kDone;
    if (!is_interior(ref)) {
        leaf = ref;
        ref = kDone;
    }
    st.sp = 0;
    for (;;) {
        while (is_interior(ref)) {
            if (wide) {
                pops += kTreeWidth / 2;
                ref = visit_wide<false, NEAR_ROWS>(S, ref, o, vinv, sel, tmin, tmax, st);
            } else {
                pops++;
                ref = visit_interior<false>(S, ref, o, vinv, tmin, tmax, __builtin_inff(), st, true);
            }
            if (is_leaf_ref(ref) && leaf == kDone) {
                leaf = ref;
                ref = st.pop_next(__builtin_inff());
            }
            if (__all(leaf != kDone))
                break;
        }
        while (leaf != kDone) {
            uint32_t s0 = leaf_start(leaf), cnt = leaf_count(leaf);
            for (uint32_t i = s0; i < s0 + cnt; i++) {
                const float4 *tp = tris + 3 * static_cast<size_t>(i);
                float4 a = ldg(tp), b = ldg(tp + 1), c = ldg(tp + 2);
                float c1, c2, t0;
                tests++;
                if (tri_hit(o, v, xyz(a), xyz(b), xyz(c), tmin, tmax, &c1, &c2, &t0))
                    return true;
            }
            if (is_leaf_ref(ref)) {
                leaf = ref;
                ref = st.pop_next(__builtin_inff());
            } else {
                leaf = kDone;
            }
        }
        if (ref == kDone)
            break;
    }
    return false;
}

// Hit record of pathspace.cpp:385-409 plus the material / light ids of the hit mesh.
struct HitRec {
    f3 p, n;
    float t;
    int32_t geo, mat, light;
};

// Hit record from the hit triangle's positions (a, b, c) and shading attributes (na, nb, nc).
X3D_DI HitRec finish_hit(ClosestHit const &h, float4 a, float4 b, float4 c, float4 na, float4 nb, float4 nc) {
    float b0 = 1.0f - h.b1 - h.b2;
    HitRec r;
    r.p = b0 * xyz(a) + h.b1 * xyz(b) + h.b2 * xyz(c);
    r.n = normalize(b0 * xyz(na) + h.b1 * xyz(nb) + h.b2 * xyz(nc));
    r.t = h.t;
    r.geo = static_cast<int32_t>(__float_as_uint(a.w));
    r.mat = static_cast<int32_t>(__float_as_uint(na.w));
    r.light = static_cast<int32_t>(__float_as_uint(nb.w));
    return r;
}

X3D_DI HitRec finish_hit(DevScene const &S, ClosestHit const &h) {
    const float4 *tris = static_cast<const float4 *>(S.tris);
    const float4 *attr = static_cast<const float4 *>(S.tri_attr);
    uint32_t i = static_cast<uint32_t>(h.prim);
    return finish_hit(h, tris[3 * i], tris[3 * i + 1], tris[3 * i + 2], attr[3 * i], attr[3 * i + 1], attr[3 * i + 2]);
}

} // namespace x3d
