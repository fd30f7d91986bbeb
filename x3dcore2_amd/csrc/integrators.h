// Per-sample estimators executed by the render kernels. Each function consumes the PCG stream
// of one (seed, pixel, sample) and returns that sample's radiance; the float operation order is
// the one of the CPU oracle, so the film is reproducible bit for bit.
#pragma once

#include "traverse.h"

namespace x3d {

constexpr uint32_t kRefModeMaxDepth = 64; // RR p = 0.5 from depth 2: P(depth >= 64) = 2^-62

struct LightSample {
    int32_t light;
    f3 p, n;
    float area_dens; // area density times light selection probability
    bool sky;
};

// basic_light_sources::sample_light (lightsources.cpp:59-75), then a point on the chosen
// light: on an area light a triangle by area, sqrt-barycentric point and interpolated normal
// (geometry.cpp:44-66); on the sky sky_light::sample_emssion_surface (light.cpp:117-129):
// u cosine-distributed about +z, p = (u + ref) * dia, n = -u, density (z . u) / pi. Draws
// consumed: 4 (the sky skips the second).
X3D_DI LightSample sample_light(DevScene const &S, Pcg &g) {
    LightSample ls;
    float e = g.draw() * S.total_power;
    uint32_t lo = 0, hi = S.num_lights;
    while (lo < hi) {
        uint32_t mi = (lo + hi) >> 1;
        if (S.light_cum_power[mi] < e)
            lo = mi + 1;
        else
            hi = mi;
    }
    if (lo >= S.num_lights)
        lo = S.num_lights - 1;
    DevLight const L = S.lights[lo];
    ls.light = static_cast<int32_t>(lo);
    ls.sky = L.kind == kLightSky;
    if (ls.sky) {
        (void)g.draw();
        float e0 = g.draw();
        float e1 = g.draw();
        const f3 z = mk(0.0f, 0.0f, 1.0f);
        f3 u = cos_hemisphere(z, e0, e1);
        ls.n = -u;
        ls.p = (u + mk(L.sky_ref[0], L.sky_ref[1], L.sky_ref[2])) * L.sky_dia;
        ls.area_dens = (dot(z, u) / kPi) * L.prob;
        return ls;
    }
    float ea = g.draw() * L.area;
    uint32_t a = 0, b = L.tri_count;
    const float *cum = S.light_tri_cum + L.tri_begin;
    while (a < b) {
        uint32_t mi = (a + b) >> 1;
        if (cum[mi] < ea)
            a = mi + 1;
        else
            b = mi;
    }
    if (a >= L.tri_count)
        a = L.tri_count - 1;
    float u = g.draw();
    float v = g.draw();
    float r = __builtin_sqrtf(u);
    float b0 = 1.0f - r;
    float b1 = r * v;
    float b2 = 1.0f - b0 - b1;
    const float4 *lt = static_cast<const float4 *>(S.light_tris) + 6u * (L.tri_begin + a);
    ls.p = b0 * xyz(lt[0]) + b1 * xyz(lt[1]) + b2 * xyz(lt[2]);
    ls.n = normalize(b0 * xyz(lt[3]) + b1 * xyz(lt[4]) + b2 * xyz(lt[5]));
    ls.area_dens = (1.0f / L.area) * L.prob;
    return ls;
}

X3D_DI f3 light_rad(DevScene const &S, int32_t l) {
    DevLight const &L = S.lights[l];
    return mk(L.rad[0], L.rad[1], L.rad[2]);
}

// Radiance a light sample sends to point xp of normal xn, before visibility: area_light::eval
// (light.cpp:41-52, L cos_i cos_o / r^2 when both cosines are positive) or sky_light::eval
// (light.cpp:131-137, L max(cos_o, 0) max(cos_i, 0), no fall-off). False when the result is
// within 1e-5 of zero (transport_illum_source, pathtracer.cpp:160-162). l = xp - ls.p and, for
// an area light, r2 = |l|^2 out.
X3D_DI bool light_arrives(DevScene const &S, LightSample const &ls, f3 xp, f3 xn, f3 &illum, f3 &l, float &r2) {
    l = xp - ls.p;
    if (ls.sky) {
        f3 in = normalize(l);
        float co = dot(ls.n, in);
        float ci = dot(xn, -in);
        r2 = 0.0f;
        illum = (light_rad(S, ls.light) * smax(co, 0.0f)) * smax(ci, 0.0f);
        return !near0(illum);
    }
    r2 = dot(l, l);
    f3 in = l / __builtin_sqrtf(r2);
    float co = dot(ls.n, in);
    float ci = dot(xn, -in);
    if (!(co > 0 && ci > 0))
        return false;
    illum = ((light_rad(S, ls.light) * ci) * co) / r2;
    return !near0(illum);
}

// Emission direction of a light sample and its solid-angle density: cosine-distributed about
// the area light's normal (area_light::sample_emssion, light.cpp:26-32), or the sky's inward
// normal -u with density 1 (sky_light::sample_emssion, light.cpp:102-115).
X3D_DI f3 emission_dir(LightSample const &em, float e0, float e1, float &w_dens) {
    if (em.sky) {
        w_dens = 1.0f;
        return em.n;
    }
    f3 w = cos_hemisphere(em.n, e0, e1);
    w_dens = dot(em.n, w) / kPi;
    return w;
}

// area_light::radiance (light.cpp:62-71): one-sided.
X3D_DI f3 light_radiance(DevScene const &S, int32_t l, f3 w, f3 n) {
    return dot(n, w) > 0.0f ? light_rad(S, l) : mk(0.0f, 0.0f, 0.0f);
}

// Subpath vertex of the bidirectional tracer (sampled_pathlet, pathtracer.cpp:12-45): position,
// shading normal, unit vector towards the previous vertex, prefix transport, material.
constexpr uint32_t kMaxPathLen = 16; // upper bound of the per-subpath vertex count
struct Pathlet {
    f3 p, n, v, transport;
    int32_t mat;
};

// One lane's subpath in global memory: four float4 rows per vertex ((p, mat), n, v, transport),
// in one contiguous block per wave (the 64 lanes' rows of one (vertex, field) pair side by side),
// so the subpaths take no private (scratch) memory.
constexpr uint32_t kPathletRows = 4;
struct PathletRows {
    float4 *base; // the wave's block (wave-uniform); lane l's rows at l, l + 64, ...
    X3D_DI float4 &row(uint32_t k, uint32_t f) const { return base[(kPathletRows * k + f) * 64u + __lane_id()]; }
    X3D_DI Pathlet load(uint32_t k) const {
        float4 a = row(k, 0), b = row(k, 1), c = row(k, 2), d = row(k, 3);
        return Pathlet{xyz(a), xyz(b), xyz(c), xyz(d), static_cast<int32_t>(__float_as_uint(a.w))};
    }
    X3D_DI void store(uint32_t k, Pathlet const &v) const {
        row(k, 0) = make_float4(v.p.x, v.p.y, v.p.z, __uint_as_float(static_cast<uint32_t>(v.mat)));
        row(k, 1) = make_float4(v.n.x, v.n.y, v.n.z, 0.0f);
        row(k, 2) = make_float4(v.v.x, v.v.y, v.v.z, 0.0f);
        row(k, 3) = make_float4(v.transport.x, v.transport.y, v.transport.z, 0.0f);
    }
};

// Record of a replayed MIS path (for the adjoint), per vertex: shading normal and material,
// outgoing direction and flags (kAdjNee: NEE contributed; kAdjCont: the BSDF sample continued the
// path), sampled BSDF direction, light-sample direction, the NEE contribution made at the vertex
// and the emission reached through the BSDF sample (both weighted by throughput and MIS weight).
// The rows live in global memory, one float4 each, in one contiguous block per wave (the 64
// lanes' rows of one (vertex, field) pair side by side), so the records take no private
// (scratch) memory and a wave's record stays within a few pages.
constexpr uint32_t kAdjNee = 1u, kAdjCont = 2u;
constexpr uint32_t kAdjMaxDepth = 8; // deepest path the adjoint records
constexpr uint32_t kAdjRows = 6;     // float4 rows per recorded vertex
struct AdjRecord {
    float4 *base; // the wave's block (wave-uniform); lane l's rows at l, l + 64, ...
    X3D_DI float4 &row(uint32_t k, uint32_t f) const { return base[(kAdjRows * k + f) * 64u + __lane_id()]; }
    // the record block of the calling lane's wave in `buf`, given the wave's global index
    X3D_DI static AdjRecord of_lane(float4 *buf, uint32_t wave) {
        return AdjRecord{buf + static_cast<size_t>(__builtin_amdgcn_readfirstlane(wave)) * (kAdjRows * kAdjMaxDepth * 64u)};
    }
};

// State of one MIS path between bounces (see Tracer::mis_begin / mis_bounce).
struct PathState {
    HitRec x;
    f3 wo, beta, L;
    Pcg g;
    uint32_t k;
    bool alive;
};

// NEAR_ROWS: whether traversal keeps the ray's near-plane row offsets in registers (fewer VALU
// instructions per wide visit) or takes per-axis min / max (three registers fewer).
template <int WG, int LDSD, bool NEAR_ROWS = true> struct Tracer {
    DevScene const &S;
    TravStack<WG, LDSD> &st;
    uint32_t pops = 0, tests = 0, spops = 0, stests = 0, nclosest = 0, nany = 0;

    X3D_DI bool trace(f3 o, f3 v, HitRec &out) {
        nclosest++;
        ClosestHit h = closest_hit<WG, LDSD, false, NEAR_ROWS>(S, o, v, st, pops, tests);
        if (h.prim < 0)
            return false;
        out = finish_hit(S, h);
        return true;
    }
    X3D_DI bool occluded(f3 o, f3 v, float tmin, float tmax) {
        nany++;
        return any_hit<WG, LDSD, NEAR_ROWS>(S, o, v, tmin, tmax, st, spops, stests);
    }

    // transport_illum_source (pathtracer.cpp:152-176) divided by the light sample's area
    // density (pathtracer.cpp:236-239). On success returns true with the unweighted estimate,
    // the direction to the light, the squared distance and the cosine at the light.
    X3D_DI bool direct(LightSample const &ls, f3 xp, f3 xn, int32_t xmat, f3 wo, f3 &contrib, f3 &wi, float &r2,
                       float &cos_l) {
        f3 l, illum;
        if (!light_arrives(S, ls, xp, xn, illum, l, r2))
            return false;
        float dist = norm(l);
        wi = -l / dist;
        if (occluded(xp, wi, 1e-4f, dist - 1e-3f))
            return false;
        f3 f = bsdf_eval(S.mats, xmat, xn, wo, wi);
        contrib = mul(illum, f) / ls.area_dens;
        cos_l = dot(ls.n, -wi);
        return true;
    }
    X3D_DI bool direct(LightSample const &ls, HitRec const &x, f3 wo, f3 &contrib, f3 &wi, float &r2,
                       float &cos_l) {
        return direct(ls, x.p, x.n, x.mat, wo, contrib, wi, r2, cos_l);
    }

    // Fixed-depth unidirectional path tracing: next-event estimation and BSDF sampling combined
    // with the power heuristic (Veach 1997); up to max_depth bounces after the first hit.
    // mis_begin starts the path of sample `s` at the cached first hit; each mis_bounce adds one
    // vertex (light sample + shadow ray, BSDF sample + closest hit, MIS-weighted emission).
    X3D_DI void mis_begin(PathState &ps, HitRec const &first, bool valid, f3 ray_v, Pcg g, uint32_t max_depth) {
        ps.g = g;
        ps.L = mk(0.0f, 0.0f, 0.0f);
        ps.beta = mk(1.0f, 1.0f, 1.0f);
        ps.k = 0;
        ps.alive = valid && max_depth > 0;
        if (!valid)
            return;
        ps.x = first;
        ps.wo = -ray_v;
        if (first.light >= 0)
            ps.L = ps.L + light_radiance(S, first.light, ps.wo, first.n);
    }

    // With REC, the vertex is recorded as vertex ps.k of `rec` (the adjoint's replay).
    template <bool REC = false>
    X3D_DI void mis_bounce(PathState &ps, uint32_t max_depth, AdjRecord const *rec = nullptr) {
        HitRec const &x = ps.x;
        // rows are written as soon as their values are known (nothing extra stays live across
        // the traversal calls)
        uint32_t flags = 0;
        const uint32_t k = ps.k;
        auto put = [&](uint32_t f, f3 v, float w) {
            if constexpr (REC)
                rec->row(k, f) = make_float4(v.x, v.y, v.z, w);
        };
        const f3 zero = mk(0.0f, 0.0f, 0.0f);
        put(0, x.n, __uint_as_float(static_cast<uint32_t>(x.mat)));
        put(4, zero, 0.0f);
        put(5, zero, 0.0f);
        if (S.num_lights > 0) {
            LightSample ls = sample_light(S, ps.g);
            f3 d, wi;
            float r2, cl;
            if (direct(ls, x, ps.wo, d, wi, r2, cl)) {
                // a BSDF sample never reaches the sky (it has no surface): weight 1
                float w = 1.0f;
                if (!ls.sky) {
                    float pl = ls.area_dens * r2 / cl;
                    float pb = bsdf_pdf(S.mats, x.mat, x.n, ps.wo, wi);
                    w = pl * pl / (pl * pl + pb * pb);
                }
                f3 add = mul(ps.beta, d * w);
                ps.L = ps.L + add;
                flags |= kAdjNee;
                put(3, wi, 0.0f);
                put(4, add, 0.0f);
            }
        }
        float dens;
        f3 wi = bsdf_sample(S.mats, x.mat, x.n, ps.wo, ps.g, &dens, false);
        HitRec nx;
        if (!(dens > 0.0f) || !trace(x.p, wi, nx) || dot(nx.n, -wi) <= 0.0f) {
            put(1, ps.wo, __uint_as_float(flags));
            ps.alive = false;
            return;
        }
        float cw = dot(x.n, wi);
        f3 f = bsdf_eval(S.mats, x.mat, x.n, ps.wo, wi);
        ps.beta = mul(ps.beta, f) * cw / dens;
        flags |= kAdjCont;
        put(1, ps.wo, __uint_as_float(flags));
        put(2, wi, 0.0f);
        if (nx.light >= 0) {
            float cl = dot(nx.n, -wi);
            DevLight const &Ln = S.lights[nx.light];
            float pl = (Ln.prob / Ln.area) * (nx.t * nx.t) / cl;
            float w = dens * dens / (dens * dens + pl * pl);
            f3 add = mul(ps.beta, light_radiance(S, nx.light, -wi, nx.n) * w);
            ps.L = ps.L + add;
            put(5, add, 0.0f);
        }
        ps.x = nx;
        ps.wo = -wi;
        ps.k++;
        if (ps.k >= max_depth)
            ps.alive = false;
    }

    X3D_DI f3 li_mis(HitRec const &first, bool valid, f3 ray_v, Pcg &g, uint32_t max_depth) {
        PathState ps;
        mis_begin(ps, first, valid, ray_v, g, max_depth);
        while (ps.alive)
            mis_bounce(ps, max_depth);
        return ps.L;
    }

    // unidirect_lt1_path_tracer (pathtracer.cpp:587-653): NEE without MIS, Russian roulette
    // p = 0.5 from depth 2, emission only at the first hit; recursion evaluated front to back.
    X3D_DI f3 li_lt1(HitRec const &first, bool valid, f3 ray_v, Pcg &g) {
        f3 L = mk(0.0f, 0.0f, 0.0f), beta = mk(1.0f, 1.0f, 1.0f);
        if (!valid)
            return L;
        HitRec x = first;
        f3 wo = -ray_v;
        for (uint32_t depth = 0; depth < kRefModeMaxDepth; depth++) {
            if (depth >= 2) {
                if (g.draw() >= 0.5f)
                    break;
                beta = beta / 0.5f;
            }
            if (S.num_lights > 0) {
                LightSample ls = sample_light(S, g);
                f3 d = mk(0.0f, 0.0f, 0.0f), wi;
                float r2, cl;
                if (!direct(ls, x, wo, d, wi, r2, cl))
                    d = mk(0.0f, 0.0f, 0.0f);
                L = L + mul(beta, d);
            }
            float dens;
            f3 wi = bsdf_sample(S.mats, x.mat, x.n, wo, g, &dens, true);
            if (dens == 0.0f)
                break;
            HitRec nx;
            if (!trace(x.p, wi, nx) || dot(nx.n, -wi) <= 0.0f)
                break;
            float cw = dot(x.n, wi);
            f3 f = bsdf_eval(S.mats, x.mat, x.n, wo, wi);
            beta = mul(beta, f) * cw / dens;
            x = nx;
            wo = -wi;
        }
        if (first.light >= 0)
            L = L + light_radiance(S, first.light, -ray_v, first.n);
        return L;
    }

    // unidirect_path_tracer (pathtracer.cpp:528-585): BSDF sampling only, emission at every
    // vertex, Russian roulette p = 0.5 from depth 2.
    X3D_DI f3 li_unidirect(HitRec const &first, bool valid, f3 ray_v, Pcg &g) {
        f3 L = mk(0.0f, 0.0f, 0.0f), beta = mk(1.0f, 1.0f, 1.0f);
        if (!valid)
            return L;
        HitRec x = first;
        f3 wo = -ray_v;
        for (uint32_t depth = 0; depth < kRefModeMaxDepth; depth++) {
            if (depth >= 2) {
                if (g.draw() >= 0.5f)
                    break;
                beta = beta / 0.5f;
            }
            if (x.light >= 0)
                L = L + mul(beta, light_radiance(S, x.light, wo, x.n));
            float dens;
            f3 wi = bsdf_sample(S.mats, x.mat, x.n, wo, g, &dens, true);
            if (dens == 0.0f)
                break;
            HitRec nx;
            if (!trace(x.p, wi, nx) || dot(nx.n, -wi) <= 0.0f)
                break;
            float cw = dot(x.n, wi);
            f3 f = bsdf_eval(S.mats, x.mat, x.n, wo, wi);
            beta = mul(beta, f) * cw / dens;
            x = nx;
            wo = -wi;
        }
        return L;
    }

    // sample_path (pathtracer.cpp:70-92): extends path[0..len) by BSDF sampling (the reference's
    // mixture densities) until max_len vertices, a density within 1e-5 of zero, a miss or a
    // back-facing hit, accumulating the prefix transport (f * cos) / dens of light_transport_info
    // (pathtracer.cpp:254-274); IMPORTANCE: projected_brdf, else projected_adjoint_brdf.
    // `cur` is path vertex len - 1 on entry.
    template <bool IMPORTANCE>
    X3D_DI uint32_t extend_walk(PathletRows const &path, Pathlet cur, uint32_t len, uint32_t max_len, Pcg &g) {
        while (len < max_len) {
            float dens;
            f3 i = bsdf_sample(S.mats, cur.mat, cur.n, cur.v, g, &dens, true);
            if (near0(dens))
                break;
            HitRec nx;
            if (!trace(cur.p, i, nx) || !(dot(nx.n, -i) > 0.0f))
                break;
            float cos_w = dot(cur.n, i);
            f3 f = IMPORTANCE ? bsdf_eval(S.mats, cur.mat, cur.n, i, cur.v) : bsdf_eval(S.mats, cur.mat, cur.n, cur.v, i);
            Pathlet nxt;
            nxt.p = nx.p;
            nxt.n = nx.n;
            nxt.v = -i;
            nxt.transport = mul(cur.transport, (f * cos_w) / dens);
            nxt.mat = nx.mat;
            path.store(len, nxt);
            cur = nxt;
            len++;
        }
        return len;
    }

    // bidirect_mis_path_tracer::sample (pathtracer.cpp:775-803): camera subpath from the cached
    // first hit, emission sample (light by power, point by area, cosine direction;
    // pathtracer.cpp:754-767, light.cpp:26-32), light subpath, then
    // transport_all_connectible_subpaths (pathtracer.cpp:351-452): per total length, the
    // average over its strategies with uniform weights; (0,0), every light_plen == 0 and every
    // connection strategy counts (occluded or not), cam_plen == 0 does not.
    X3D_DI f3 li_bdpt(HitRec const &first, bool valid, f3 ray_v, Pcg &g, uint32_t max_len, PathletRows const &cam,
                      PathletRows const &lig) {
        f3 rad = mk(0.0f, 0.0f, 0.0f);
        if (!valid || max_len == 0 || S.num_lights == 0)
            return rad;
        const Pathlet c0{first.p, first.n, -ray_v, mk(1.0f, 1.0f, 1.0f), first.mat};
        cam.store(0, c0);
        uint32_t cam_len = extend_walk<false>(cam, c0, 1, max_len, g);

        LightSample em = sample_light(S, g);
        float e0 = g.draw();
        float e1 = g.draw();
        float w_dens;
        f3 w = emission_dir(em, e0, e1, w_dens);
        uint32_t light_len = 0;
        HitRec v0;
        if (trace(em.p, w, v0) && dot(v0.n, -w) > 0.0f) {
            const Pathlet l0{v0.p, v0.n, -w, mk(1.0f, 1.0f, 1.0f), v0.mat};
            lig.store(0, l0);
            light_len = extend_walk<true>(lig, l0, 1, max_len, g);
        }
        float cos_e = dot(em.n, w);
        f3 emitted = cos_e > 0.0f ? light_rad(S, em.light) * cos_e : mk(0.0f, 0.0f, 0.0f);
        // a light of zero selection probability carries no power
        f3 light_emission = em.area_dens > 0.0f ? emitted / (w_dens * em.area_dens) : mk(0.0f, 0.0f, 0.0f);

        for (uint32_t plen = 1; plen <= cam_len + light_len + 1; plen++) {
            int cp = static_cast<int>(plen - 1 < cam_len ? plen - 1 : cam_len);
            uint32_t lp = plen - 1 - static_cast<uint32_t>(cp);
            f3 psum = mk(0.0f, 0.0f, 0.0f);
            float wsum = 0.0f;
            for (; cp >= 0 && lp <= light_len; lp++, cp--) {
                if (lp == 0 && cp == 0) {
                    if (first.light >= 0)
                        psum = psum + light_radiance(S, first.light, c0.v, c0.n);
                    wsum = wsum + 1.0f;
                } else if (lp == 0) {
                    Pathlet const cj = cam.load(cp - 1);
                    f3 tr = mk(0.0f, 0.0f, 0.0f), wi;
                    float r2, cl;
                    if (!direct(em, cj.p, cj.n, cj.mat, cj.v, tr, wi, r2, cl))
                        tr = mk(0.0f, 0.0f, 0.0f);
                    psum = psum + mul(tr, cj.transport);
                    wsum = wsum + 1.0f;
                } else if (cp > 0) {
                    Pathlet const lj = lig.load(lp - 1);
                    Pathlet const cj = cam.load(cp - 1);
                    f3 join = cj.p - lj.p;
                    float dist = norm(join);
                    join = join / dist;
                    float cos_wo = dot(lj.n, join);
                    float cos_wi = dot(cj.n, -join);
                    if (cos_wo > 0.0f && cos_wi > 0.0f && !occluded(lj.p, join, 1e-3f, dist - 1e-3f)) {
                        f3 sub = mul(light_emission, lj.transport);
                        float geo = cos_wi * cos_wo / (dist * dist);
                        f3 fl = bsdf_eval(S.mats, lj.mat, lj.n, join, lj.v);
                        f3 fc = bsdf_eval(S.mats, cj.mat, cj.n, cj.v, -join);
                        f3 tr = mul(mul(sub, fl), fc) * geo;
                        psum = psum + mul(tr, cj.transport);
                    }
                    wsum = wsum + 1.0f;
                }
            }
            if (wsum > 0.0f)
                rad = rad + psum / wsum;
        }
        return rad;
    }

    // bidirect_lt2_path_tracer (pathtracer.cpp:655-752), evaluated front to back. At every
    // camera vertex, join_with_light_paths: one NEE sample p1 and one light path of a single
    // vertex (emission sample, closest hit) joined to the camera vertex, p2; their mean (at the
    // first vertex p1 + p2 / 2). As in the reference, a light path that misses the scene makes
    // the vertex's term zero (p1 included), and a BSDF sample whose hit lies below the vertex's
    // surface (cos < 0) ends the path without the vertex's term. Russian roulette p = 0.5 from
    // depth 1. Where the BSDF sample has zero density the path ends with the vertex's term (the
    // reference divides by that zero).
    X3D_DI f3 li_lt2(HitRec const &first, bool valid, f3 ray_v, Pcg &g) {
        f3 L = mk(0.0f, 0.0f, 0.0f), beta = mk(1.0f, 1.0f, 1.0f);
        if (!valid)
            return L;
        HitRec x = first;
        f3 wo = -ray_v;
        for (uint32_t depth = 0; depth < kRefModeMaxDepth; depth++) {
            if (depth >= 1) {
                if (g.draw() >= 0.5f)
                    break;
                beta = beta / 0.5f;
            }
            f3 B = mk(0.0f, 0.0f, 0.0f);
            if (S.num_lights > 0) {
                LightSample ls = sample_light(S, g);
                f3 p1 = mk(0.0f, 0.0f, 0.0f), wl;
                float r2, cl;
                if (!direct(ls, x, wo, p1, wl, r2, cl))
                    p1 = mk(0.0f, 0.0f, 0.0f);
                LightSample em = sample_light(S, g);
                float e0 = g.draw();
                float e1 = g.draw();
                float w_dens;
                f3 w = emission_dir(em, e0, e1, w_dens);
                HitRec lv;
                if (trace(em.p, w, lv)) {
                    float cos_e = dot(em.n, w);
                    f3 emitted = cos_e > 0.0f ? light_rad(S, em.light) * cos_e : mk(0.0f, 0.0f, 0.0f);
                    f3 illum = em.area_dens > 0.0f ? emitted / (em.area_dens * w_dens) : mk(0.0f, 0.0f, 0.0f);
                    f3 tray = -w;
                    f3 jp = x.p - lv.p;
                    float dist = norm(jp);
                    jp = jp / dist;
                    float cos_w2 = dot(lv.n, tray);
                    float cos_wo = dot(lv.n, jp);
                    float cos_wi = dot(x.n, -jp);
                    if (cos_wo > 0.0f && cos_wi > 0.0f && cos_w2 > 0.0f && !occluded(lv.p, jp, 1e-4f, dist - 1e-3f)) {
                        f3 f2 = mul(illum, bsdf_eval(S.mats, lv.mat, lv.n, jp, tray)) * cos_w2;
                        f3 p2 = mul(f2 * cos_wo / (dist * dist), bsdf_eval(S.mats, x.mat, x.n, wo, -jp)) * cos_wi;
                        B = depth == 0 ? p1 + 0.5f * p2 : 0.5f * (p1 + p2);
                    } else {
                        B = p1;
                    }
                }
            }
            float dens;
            f3 wi = bsdf_sample(S.mats, x.mat, x.n, wo, g, &dens, true);
            if (dens == 0.0f) {
                L = L + mul(beta, B);
                break;
            }
            HitRec nx;
            if (!trace(x.p, wi, nx)) {
                L = L + mul(beta, B);
                break;
            }
            float cw = dot(x.n, wi);
            if (cw < 0.0f)
                break;
            L = L + mul(beta, B);
            f3 f = bsdf_eval(S.mats, x.mat, x.n, wo, wi);
            beta = mul(beta, f) * cw / dens;
            x = nx;
            wo = -wi;
        }
        if (first.light >= 0 && dot(first.n, -ray_v) > 0.0f)
            L = L + light_rad(S, first.light) * dot(first.n, -ray_v);
        return L;
    }

    // position_tracer (pathtracer.cpp:474-490): the first hit's position in the path space's
    // bounds, (p - min) / (max - min) per axis.
    X3D_DI f3 li_position(HitRec const &first, bool valid) const {
        if (!valid)
            return mk(0.0f, 0.0f, 0.0f);
        return mk((first.p.x - S.space_lo[0]) / S.space_range[0], (first.p.y - S.space_lo[1]) / S.space_range[1],
                  (first.p.z - S.space_lo[2]) / S.space_range[2]);
    }

    // direct_path_tracer (pathtracer.cpp:508-526).
    X3D_DI f3 li_direct(HitRec const &first, bool valid, f3 ray_v, Pcg &g) {
        f3 L = mk(0.0f, 0.0f, 0.0f);
        if (!valid)
            return L;
        f3 wo = -ray_v;
        if (S.num_lights > 0) {
            LightSample ls = sample_light(S, g);
            f3 d, wi;
            float r2, cl;
            if (direct(ls, first, wo, d, wi, r2, cl))
                L = d;
        }
        if (first.light >= 0 && dot(first.n, wo) > 0.0f)
            L = L + light_rad(S, first.light) * dot(first.n, wo);
        return L;
    }
};

} // namespace x3d
