// Test-infrastructure harness around the reference renderer (DaviesX/x3dcore2, "e8yescg").
//
// This file is compiled together with the reference's own hot-path sources, where they lie
// under the reference checkout, into oracle/_ref/libx3dref.so (see oracle/Makefile). It only
// drives reference code through its public classes; it contains no restatement of the
// reference's algorithms except the 40-line sample accumulation of pt_image_renderer::render
// (corelib/src/renderer.cpp:35-166), which cannot be compiled here because
// renderer.h pulls in Qt through frame.h (frame.h:6).
//
// Used only by tests/ and by bench.py's CPU-baseline leg. Never part of the product.
//
// Hit identification: the reference's intersect_info does not expose the hit triangle
// (pathspace.cpp:385-409 returns geo, t, vertex, normal, uv). Every mesh is therefore loaded
// de-indexed (three private vertices per triangle, same positions and normals), with the
// texture coordinate of all three vertices set to (triangle index, geometry index). The
// interpolated uv the reference returns then identifies the winning triangle exactly, while the
// BVH build (which reads only positions) and the traversal are unchanged.

#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <exception>
#include <initializer_list>
#include <iosfwd>
#include <map>
#include <memory>
#include <optional>
#include <ostream>
#include <random>
#include <set>
#include <stdint.h>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

// Read-only access to the flattened BVH, primitive array and geometry map for layout dumps, and
// to the sky light's reference point and diameter (all standard headers are included above, so
// only the reference's classes are affected).
#define private public
#define protected public
#include "pathspace.h"
#include "light.h"
#undef protected
#undef private

#include "camera.h"
#include "geometry.h"
#include "lightsources.h"
#include "material.h"
#include "materialcontainer.h"
#include "pathtracer.h"
#include "pathtracerfact.h"
#include "tensor.h"

namespace {

struct ref_scene {
    std::vector<std::shared_ptr<e8::if_material>> mats;
    std::vector<std::shared_ptr<e8::trimesh>> meshes;
    std::vector<std::shared_ptr<e8::if_light>> lights;
    std::unique_ptr<e8::bvh_path_space_layout> bvh;
    std::unique_ptr<e8::linear_path_space_layout> linear;
    std::unique_ptr<e8::default_material_container> mat_container;
    std::unique_ptr<e8::basic_light_sources> light_sources;
    std::map<e8::if_geometry const *, uint32_t> geo_index;
    bool committed = false;
};

e8util::vec3 v3(float const *p) { return e8util::vec3({p[0], p[1], p[2]}); }

void put3(float *dst, e8util::vec3 const &v) {
    dst[0] = v(0);
    dst[1] = v(1);
    dst[2] = v(2);
}

e8util::mat44 mat_from_colmajor(float const *m) {
    e8util::mat44 r;
    for (unsigned j = 0; j < 4; j++)
        for (unsigned i = 0; i < 4; i++)
            r(i, j) = m[j * 4 + i];
    return r;
}

int decode_hit(e8::intersect_info const &info, ref_scene const *s, int32_t *prim_geo,
               int32_t *prim_tri) {
    if (!info.valid()) {
        *prim_geo = -1;
        *prim_tri = -1;
        return 0;
    }
    *prim_tri = static_cast<int32_t>(std::lround(info.uv(0)));
    *prim_geo = static_cast<int32_t>(s->geo_index.at(info.geo));
    return 1;
}

} // namespace

extern "C" {

void *ref_scene_new() { return new ref_scene(); }

void ref_scene_free(void *h) { delete static_cast<ref_scene *>(h); }

int ref_add_oren_nayar(void *h, float const *albedo, float sigma) {
    ref_scene *s = static_cast<ref_scene *>(h);
    s->mats.push_back(std::make_shared<e8::oren_nayar>("on", v3(albedo), sigma));
    return static_cast<int>(s->mats.size()) - 1;
}

int ref_add_cook_torr(void *h, float const *albedo, float roughness, float ior_re, float ior_im) {
    ref_scene *s = static_cast<ref_scene *>(h);
    s->mats.push_back(std::make_shared<e8::cook_torr>("ct", v3(albedo), roughness,
                                                      std::complex<float>(ior_re, ior_im)));
    return static_cast<int>(s->mats.size()) - 1;
}

int ref_add_mixture(void *h, int m0, int m1, float ratio) {
    ref_scene *s = static_cast<ref_scene *>(h);
    s->mats.push_back(std::make_shared<e8::mat_mixture>("mix", s->mats.at(m0)->copy(),
                                                        s->mats.at(m1)->copy(), ratio));
    return static_cast<int>(s->mats.size()) - 1;
}

// verts/normals: float[3*nv]; tris: uint32[3*nt]; mat: index from ref_add_* or -1 (fail-safe).
int ref_add_mesh(void *h, float const *verts, float const *normals, uint32_t nv,
                 uint32_t const *tris, uint32_t nt, int mat) {
    (void)nv;
    ref_scene *s = static_cast<ref_scene *>(h);
    uint32_t geo = static_cast<uint32_t>(s->meshes.size());
    std::vector<e8util::vec3> vv, nn;
    std::vector<e8util::vec2> uv;
    std::vector<e8::triangle> tt;
    vv.reserve(3 * nt);
    nn.reserve(3 * nt);
    uv.reserve(3 * nt);
    tt.reserve(nt);
    for (uint32_t k = 0; k < nt; k++) {
        for (unsigned c = 0; c < 3; c++) {
            uint32_t vi = tris[3 * k + c];
            vv.push_back(v3(verts + 3 * vi));
            nn.push_back(v3(normals + 3 * vi));
            uv.push_back(e8util::vec2({static_cast<float>(k), static_cast<float>(geo)}));
        }
        tt.push_back(e8::triangle({3 * k, 3 * k + 1, 3 * k + 2}));
    }
    auto mesh = std::make_shared<e8::trimesh>("mesh");
    mesh->vertices(vv);
    mesh->normals(nn);
    mesh->texcoords(uv);
    mesh->triangles(tt);
    mesh->update();
    if (mat >= 0)
        mesh->attach_material(s->mats.at(static_cast<size_t>(mat))->id());
    s->meshes.push_back(mesh);
    return static_cast<int>(geo);
}

int ref_add_area_light(void *h, int mesh, float const *rad) {
    ref_scene *s = static_cast<ref_scene *>(h);
    s->lights.push_back(
        std::make_shared<e8::area_light>("light", s->meshes.at(static_cast<size_t>(mesh)), v3(rad)));
    return static_cast<int>(s->lights.size()) - 1;
}

// sky_light (light.cpp:87-163). Nothing in the reference calls set_scene_boundary; ref_commit
// calls it with the path space's bounds before the light sources are committed.
int ref_add_sky_light(void *h, float const *rad) {
    ref_scene *s = static_cast<ref_scene *>(h);
    s->lights.push_back(std::make_shared<e8::sky_light>(v3(rad)));
    return static_cast<int>(s->lights.size()) - 1;
}

int ref_commit(void *h) {
    ref_scene *s = static_cast<ref_scene *>(h);
    e8util::mat44 id = e8util::mat44_scale(1.0f);
    s->bvh = std::make_unique<e8::bvh_path_space_layout>();
    s->linear = std::make_unique<e8::linear_path_space_layout>();
    s->mat_container = std::make_unique<e8::default_material_container>();
    s->light_sources = std::make_unique<e8::basic_light_sources>();
    for (auto const &m : s->mats)
        s->mat_container->load(*m, id);
    s->mat_container->commit();
    for (auto const &g : s->meshes) {
        s->bvh->load(*g, id);
        s->linear->load(*g, id);
    }
    s->bvh->commit();
    s->linear->commit();
    s->geo_index.clear();
    for (auto const &p : s->bvh->m_geometries) {
        // The path space holds transformed copies; the texcoord v of any vertex is the index.
        s->geo_index[p.second.get()] = static_cast<uint32_t>(std::lround(p.second->texcoords()[0](1)));
    }
    for (auto const &p : s->linear->m_geometries) {
        s->geo_index[p.second.get()] = static_cast<uint32_t>(std::lround(p.second->texcoords()[0](1)));
    }
    for (auto const &l : s->lights) {
        l->set_scene_boundary(s->bvh->aabb());
        s->light_sources->load(*l, id);
    }
    if (!s->lights.empty())
        s->light_sources->commit();
    s->committed = true;
    return 0;
}

void ref_bvh_size(void *h, uint32_t *num_nodes, uint32_t *num_prims) {
    ref_scene *s = static_cast<ref_scene *>(h);
    *num_nodes = static_cast<uint32_t>(s->bvh->m_bvh.size());
    *num_prims = static_cast<uint32_t>(s->bvh->m_prims.size());
}

// bounds: float[6*nodes] (min xyz, max xyz); meta: uint32[4*nodes] (num_prims, split_axis,
// next_child, prim_start); prims: int32[2*prims] (geometry index, triangle index).
void ref_bvh_dump(void *h, float *bounds, uint32_t *meta, int32_t *prims) {
    ref_scene *s = static_cast<ref_scene *>(h);
    auto const &nodes = s->bvh->m_bvh;
    for (size_t n = 0; n < nodes.size(); n++) {
        put3(bounds + 6 * n, nodes[n].bound.min());
        put3(bounds + 6 * n + 3, nodes[n].bound.max());
        meta[4 * n + 0] = nodes[n].num_prims;
        meta[4 * n + 1] = nodes[n].split_axis;
        meta[4 * n + 2] = nodes[n].next_child;
        meta[4 * n + 3] = nodes[n].prim_start;
    }
    auto const &ps = s->bvh->m_prims;
    for (size_t i = 0; i < ps.size(); i++) {
        prims[2 * i + 0] = static_cast<int32_t>(ps[i].i_geo);
        prims[2 * i + 1] = static_cast<int32_t>(ps[i].tri(0) / 3);
    }
}

// rays: float[6*n] (origin, direction). Outputs per ray: geometry, triangle (-1 on miss), t,
// hit vertex, shading normal. use_linear selects the brute-force layout (pathspace.cpp:36-92).
void ref_intersect(void *h, float const *rays, uint32_t n, int use_linear, int32_t *geo,
                   int32_t *tri, float *t, float *vertex, float *normal) {
    ref_scene *s = static_cast<ref_scene *>(h);
    e8::if_path_space const &ps =
        use_linear ? static_cast<e8::if_path_space const &>(*s->linear) : *s->bvh;
    for (uint32_t k = 0; k < n; k++) {
        e8util::ray r(v3(rays + 6 * k), v3(rays + 6 * k + 3));
        e8::intersect_info info = ps.intersect(r);
        if (decode_hit(info, s, geo + k, tri + k)) {
            t[k] = info.t;
            put3(vertex + 3 * k, info.vertex);
            put3(normal + 3 * k, info.normal);
        } else {
            t[k] = INFINITY;
            put3(vertex + 3 * k, e8util::vec3());
            put3(normal + 3 * k, e8util::vec3());
        }
    }
}

void ref_occluded(void *h, float const *rays, float const *tmin, float const *tmax, uint32_t n,
                  int use_linear, uint8_t *out) {
    ref_scene *s = static_cast<ref_scene *>(h);
    e8::if_path_space const &ps =
        use_linear ? static_cast<e8::if_path_space const &>(*s->linear) : *s->bvh;
    for (uint32_t k = 0; k < n; k++) {
        e8util::ray r(v3(rays + 6 * k), v3(rays + 6 * k + 3));
        float t;
        out[k] = ps.has_intersect(r, tmin[k], tmax[k], t) ? 1 : 0;
    }
}

// BSDF eval f(o, i) at shading normal n for material index mat (tst_material.cpp:81 convention).
void ref_bsdf_eval(void *h, int mat, uint32_t n, float const *nrm, float const *o, float const *i,
                   float *f) {
    ref_scene *s = static_cast<ref_scene *>(h);
    e8::if_material const &m = *s->mats.at(static_cast<size_t>(mat));
    for (uint32_t k = 0; k < n; k++) {
        e8util::vec3 r = m.eval(e8util::vec2(), v3(nrm + 3 * k), v3(o + 3 * k), v3(i + 3 * k));
        put3(f + 3 * k, r);
    }
}

// BSDF sample with the reference's rng (minstd_rand0 + uniform_real_distribution<float>) seeded
// once and drawn sequentially, as tst_material.cpp:47-52 does.
void ref_bsdf_sample(void *h, int mat, uint32_t seed, uint32_t n, float const *nrm, float const *o,
                     float *i_out, float *dens_out) {
    ref_scene *s = static_cast<ref_scene *>(h);
    e8::if_material const &m = *s->mats.at(static_cast<size_t>(mat));
    e8util::rng rng(seed);
    for (uint32_t k = 0; k < n; k++) {
        float dens = -1.0f;
        e8util::vec3 r = m.sample(&rng, &dens, e8util::vec2(), v3(nrm + 3 * k), v3(o + 3 * k));
        put3(i_out + 3 * k, r);
        dens_out[k] = dens;
    }
}

// Draws from the reference rng, for pinning the oracle's copy of its uniform stream.
void ref_rng_draws(uint32_t seed, uint32_t n, float *out) {
    e8util::rng rng(seed);
    for (uint32_t k = 0; k < n; k++)
        out[k] = rng.draw();
}

// Pinhole camera rays (camera.cpp:38-48) after pinhole_camera::transform (camera.cpp:63-71).
// trans: column-major 4x4, element (i, j) at trans[4*j + i].
void ref_camera_rays(float sensor, float focal, float aspect, float const *trans, uint32_t w,
                     uint32_t h, float *rays) {
    e8::pinhole_camera cam(sensor, focal, aspect);
    std::unique_ptr<e8::if_camera> tcam = cam.transform(mat_from_colmajor(trans));
    e8util::rng rng(1361);
    for (uint32_t j = 0; j < h; j++) {
        for (uint32_t i = 0; i < w; i++) {
            float pdf;
            e8util::ray r = tcam->sample(rng, i, j, w, h, pdf);
            put3(rays + 6 * (i + j * w), r.o());
            put3(rays + 6 * (i + j * w) + 3, r.v());
        }
    }
}

// Renders with reference tracer `pt_type` (pathtracerfact.h:12 order) using `threads` sampling
// tasks, each seeded i*1361+33 and drawing ceil(spp/threads) samples (renderer.cpp:96-104, 124-132).
// rgba: float[4*w*h], row 0 = top. Returns the number of samples actually taken per pixel.
uint32_t ref_render(void *h, int pt_type, float sensor, float focal, float aspect,
                    float const *trans, uint32_t w, uint32_t hh, uint32_t spp, uint32_t threads,
                    float *rgba) {
    ref_scene *s = static_cast<ref_scene *>(h);
    e8::pinhole_camera cam(sensor, focal, aspect);
    std::unique_ptr<e8::if_camera> tcam = cam.transform(mat_from_colmajor(trans));
    e8util::rng cam_rng(1361);
    std::vector<e8util::ray> rays(static_cast<size_t>(w) * hh);
    for (uint32_t j = 0; j < hh; j++)
        for (uint32_t i = 0; i < w; i++) {
            float pdf;
            rays[i + j * w] = tcam->sample(cam_rng, i, j, w, hh, pdf);
        }
    e8::if_path_tracer::first_hits first_hits =
        e8::if_path_tracer::compute_first_hit(rays, *s->bvh, *s->light_sources);

    if (threads == 0)
        threads = 1;
    uint32_t allocated =
        static_cast<uint32_t>(std::ceil(static_cast<float>(spp) / static_cast<float>(threads)));
    std::vector<std::vector<e8util::vec3>> estimates(threads);
    std::vector<std::thread> pool;
    e8::pathtracer_factory fact(static_cast<e8::pathtracer_factory::pt_type>(pt_type),
                                e8::pathtracer_factory::options());
    for (uint32_t t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            std::unique_ptr<e8::if_path_tracer> pt(fact.create());
            e8util::rng rng(t * 1361 + 33);
            std::vector<e8util::vec3> &acc = estimates[t];
            acc.assign(rays.size(), e8util::vec3());
            for (uint32_t k = 0; k < allocated; k++) {
                std::vector<e8util::vec3> est =
                    pt->sample(rng, rays, first_hits, *s->bvh, *s->mat_container, *s->light_sources);
                for (size_t p = 0; p < est.size(); p++)
                    acc[p] += est[p];
            }
        });
    }
    for (auto &th : pool)
        th.join();
    std::vector<e8util::vec4> film(rays.size(), e8util::vec4());
    for (uint32_t t = 0; t < threads; t++)
        for (size_t p = 0; p < rays.size(); p++)
            film[p] += estimates[t][p].homo(1.0f);
    float scale = 1.0f / (allocated * threads);
    for (size_t p = 0; p < rays.size(); p++) {
        film[p] *= scale;
        for (unsigned c = 0; c < 4; c++)
            rgba[4 * p + c] = film[p](c);
    }
    return allocated * threads;
}

// Light state after commit: power() (3), and for a sky light m_ref_p (3) and m_dia; zeros
// for an area light.
void ref_light_info(void *h, int light, float *out) {
    ref_scene *s = static_cast<ref_scene *>(h);
    e8::if_light const &L = *s->lights.at(static_cast<size_t>(light));
    put3(out, L.power());
    out[3] = out[4] = out[5] = out[6] = 0.0f;
    if (auto const *sky = dynamic_cast<e8::sky_light const *>(&L)) {
        put3(out + 3, sky->m_ref_p);
        out[6] = sky->m_dia;
    }
}

// if_light::eval (light.cpp:41-52, 131-137) of light `light`: l float[3n] (target - light point),
// nl / nt the light and target normals -> out float[3n].
void ref_light_eval(void *h, int light, uint32_t n, float const *l, float const *nl, float const *nt,
                    float *out) {
    ref_scene *s = static_cast<ref_scene *>(h);
    e8::if_light const &L = *s->lights.at(static_cast<size_t>(light));
    for (uint32_t k = 0; k < n; k++)
        put3(out + 3 * k, L.eval(v3(l + 3 * k), v3(nl + 3 * k), v3(nt + 3 * k)));
}

// trimesh::transform (geometry.cpp:76-87) of a mesh with the given column-major 4x4 matrix
// (e8util::mat44's initializer order): out_verts / out_normals float[3*nv]. Returns 0, or -1
// when the reference throws (matrix not invertible).
int ref_transform_mesh(float const *trans, uint32_t nv, float const *verts, float const *normals,
                       float *out_verts, float *out_normals) {
    e8util::mat44 m({trans[0], trans[1], trans[2], trans[3], trans[4], trans[5], trans[6], trans[7], trans[8],
                     trans[9], trans[10], trans[11], trans[12], trans[13], trans[14], trans[15]});
    e8::trimesh mesh("mesh");
    std::vector<e8util::vec3> vv, nn;
    for (uint32_t i = 0; i < nv; i++) {
        vv.push_back(v3(verts + 3 * i));
        nn.push_back(v3(normals + 3 * i));
    }
    mesh.vertices(vv);
    mesh.normals(nn);
    try {
        std::unique_ptr<e8::if_geometry> t = mesh.transform(m);
        for (uint32_t i = 0; i < nv; i++)
            for (unsigned c = 0; c < 3; c++) {
                out_verts[3 * i + c] = t->vertices()[i](c);
                out_normals[3 * i + c] = t->normals()[i](c);
            }
    } catch (std::string const &) {
        return -1;
    }
    return 0;
}

} // extern "C"
