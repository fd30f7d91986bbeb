// ORACLE — test infrastructure only. Runs host code under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by `make -C oracle sanitize`, run by tests/test_sanitizers.py):
//   - the library's Wavefront OBJ reader (x3dcore2_amd/csrc/objfile.cpp) on the reference's
//     Cornell box files and on malformed files;
//   - the library's host scene code (x3dcore2_amd/csrc/scene.cpp): the reference-layout BVH
//     build, the 4-wide tree collapse and its consistency check, the device packing, objects by
//     id with transforms (without a GPU the commit stops at its first device allocation, after
//     all host work);
//   - the oracle's scene build (BVH, light CDFs), every tracer, the firefly filter, ray queries
//     and the compositors, on the Cornell box loaded through that reader.
// Any sanitizer report aborts the process (-fno-sanitize-recover=all); a failed check returns 1.
//
//   sanitize_check CORNELL_DIR SCRATCH_DIR

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../x3dcore2_amd/csrc/objfile.h"
#include "../x3dcore2_amd/csrc/scene.h"

extern "C" {
void *orc_scene_new();
void orc_scene_free(void *h);
int orc_add_oren_nayar(void *h, float const *albedo, float sigma);
int orc_add_cook_torr(void *h, float const *albedo, float roughness, float ior_re, float ior_im);
int orc_add_mixture(void *h, int m0, int m1, float ratio);
int orc_add_mesh(void *h, float const *verts, float const *normals, uint32_t nv, uint32_t const *tris, uint32_t nt,
                 int mat);
int orc_add_area_light(void *h, int mesh, float const *rad);
int orc_commit(void *h);
void orc_bvh_size(void *h, uint32_t *nn, uint32_t *np);
void orc_intersect(void *h, float const *rays, uint32_t n, int mode, int32_t *geo, int32_t *tri, int32_t *prim,
                   float *t, float *b, float *vertex, float *normal, uint64_t *stats);
void orc_render(void *h, float const *cam, int tracer, uint32_t max_depth, uint32_t max_path_len, uint64_t seed,
                int traversal, uint32_t w, uint32_t hh, uint32_t spp, uint32_t const *pixels, uint32_t npix, int prec,
                uint32_t threads, void *rgba, uint64_t *stats);
void orc_render_firefly(void *h, float const *cam, int tracer, uint32_t max_depth, uint32_t max_path_len,
                        uint64_t seed, uint32_t w, uint32_t hh, uint32_t spp, uint32_t threads, float *rgba);
void orc_compose(float const *rgba, uint32_t w, uint32_t h, int kind, int auto_exposure, float exposure,
                 uint8_t *out, float *exposure_used, uint32_t *counted);
void orc_compose_libm(float const *rgba, uint32_t w, uint32_t h, int kind, int auto_exposure, float exposure,
                      uint8_t *out, float *exposure_used, uint32_t *counted);
void orc_compose_math(int fn, float const *x, uint32_t n, float *out);
}

namespace {

int failures = 0;

void expect(bool ok, std::string const &what) {
    if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what.c_str());
        failures++;
    }
}

std::string write_file(std::string const &dir, std::string const &name, std::string const &text) {
    std::string path = dir + "/" + name;
    std::ofstream(path) << text;
    return path;
}

} // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s CORNELL_DIR SCRATCH_DIR\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1], scratch = argv[2];
    const char *parts[] = {"left_wall", "right_wall", "back_wall", "ceiling",
                           "floor",     "left_sphere", "right_sphere", "light"};
    const int mat_of[] = {1, 2, 0, 0, 0, 3, 0, 4};

    // OBJ reader: good files, then malformed ones (each must fail with a message, not crash).
    std::vector<x3d::ObjMesh> meshes(8);
    for (int k = 0; k < 8; k++) {
        std::string e = x3d::load_obj((dir + "/" + parts[k] + ".obj").c_str(), meshes[k]);
        expect(e.empty() && !meshes[k].tris.empty(), std::string("load ") + parts[k] + ": " + e);
    }
    const char *bad[] = {
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/1/1 2/1/1 3/1/1 4/1/1\n", "v 0 0 0\nf 1 2 3\n",
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 4294967296// 1// 2//\n",   "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1/99/ 2// 3//\n",
        "v 1e99999 0 0\n",                                         "vt 0.5\nvn 1 2\nf\n",
        "v\nv \nvt\nf 1/2/3/4 5 6\n",                              "",
    };
    for (size_t k = 0; k < sizeof bad / sizeof bad[0]; k++) {
        x3d::ObjMesh m;
        std::string e = x3d::load_obj(write_file(scratch, "bad" + std::to_string(k) + ".obj", bad[k]).c_str(), m);
        expect(!e.empty(), "malformed file " + std::to_string(k) + " accepted");
    }
    {
        x3d::ObjMesh m;
        expect(!x3d::load_obj((scratch + "/absent.obj").c_str(), m).empty(), "absent file accepted");
    }

    // The library's host scene: indexed, then by id with a transform and a second commit.
    {
        x3d::Scene sc;
        for (int k = 0; k < 8; k++) {
            x3d::HostMesh hm;
            x3d::ObjMesh const &m = meshes[k];
            for (size_t v = 0; v < m.verts.size() / 3; v++) {
                hm.verts.push_back({m.verts[3 * v], m.verts[3 * v + 1], m.verts[3 * v + 2]});
                hm.normals.push_back({m.normals[3 * v], m.normals[3 * v + 1], m.normals[3 * v + 2]});
            }
            hm.tris = m.tris;
            sc.grow_space(hm.verts);
            sc.meshes.push_back(hm);
        }
        sc.lights.push_back({7, {13.665f, 9.9f, 5.175f}});
        std::string e = sc.commit();
        expect(e.empty() || e.find("hip") != std::string::npos, "indexed commit: " + e);
        expect(!sc.flat.empty() && sc.prims.size() > 1000 && sc.num_nodes4 > 0, "indexed scene: no BVH");

        x3d::Scene kd;
        kd.keyed = true;
        const float t[16] = {0.8f, 0.0f, -0.1f, 0.0f, 0.0f, 1.1f, 0.0f, 0.0f,
                             0.1f, 0.0f, 0.9f,  0.0f, 0.05f, -0.02f, 0.1f, 1.0f};
        for (int k = 0; k < 8; k++) {
            x3d::KeyedGeometry g;
            g.mesh = sc.meshes[k];
            if (k == 5)
                expect(x3d::transform_mesh(t, g.mesh.verts, g.mesh.normals).empty(), "transform");
            g.has_material = k < 4;
            g.material = 100 + k;
            kd.grow_space(g.mesh.verts);
            kd.k_geos.insert({static_cast<uint64_t>(900 - 7 * k), g});
        }
        kd.k_lights.insert({1u, x3d::KeyedLight{900 - 7 * 7, {1.0f, 1.0f, 1.0f}}});
        e = kd.commit();
        expect(e.empty() || e.find("hip") != std::string::npos, "keyed commit: " + e);
        kd.k_geos.erase(900 - 7 * 3);
        e = kd.commit();
        expect(e.empty() || e.find("hip") != std::string::npos, "keyed recommit: " + e);
        expect(kd.meshes.size() == 7 && kd.geo_ids.size() == 7, "keyed scene after unload");
        kd.k_lights.insert({2u, x3d::KeyedLight{12345, {1.0f, 1.0f, 1.0f}}});
        expect(!kd.commit().empty(), "a light on an absent geometry was accepted");
    }

    // Oracle scene from the reader's meshes.
    void *h = orc_scene_new();
    const float white[3] = {0.725f, 0.710f, 0.680f}, red[3] = {0.630f, 0.065f, 0.050f},
                green[3] = {0.140f, 0.450f, 0.091f}, glossy[3] = {0.95f, 0.95f, 0.95f}, black[3] = {0, 0, 0};
    int ids[5] = {orc_add_oren_nayar(h, white, 0.078f), orc_add_oren_nayar(h, red, 0.078f),
                  orc_add_oren_nayar(h, green, 0.078f), orc_add_cook_torr(h, glossy, 0.2f, 2.93f, 3.0f),
                  orc_add_oren_nayar(h, black, 0.078f)};
    orc_add_mixture(h, ids[0], ids[3], 0.5f);
    int light_mesh = -1;
    for (int k = 0; k < 8; k++) {
        x3d::ObjMesh const &m = meshes[k];
        light_mesh = orc_add_mesh(h, m.verts.data(), m.normals.data(), static_cast<uint32_t>(m.verts.size() / 3),
                                  m.tris.data(), static_cast<uint32_t>(m.tris.size() / 3), ids[mat_of[k]]);
    }
    const float rad[3] = {0.911f * 15.0f, 0.660f * 15.0f, 0.345f * 15.0f};
    orc_add_area_light(h, light_mesh, rad);
    orc_commit(h);
    uint32_t nn = 0, np = 0;
    orc_bvh_size(h, &nn, &np);
    expect(nn > 0 && np > 0, "empty BVH");

    float cam[19] = {0.032f, 0.035f, 4.0f / 3.0f};
    for (int k = 0; k < 16; k++)
        cam[3 + k] = k % 5 == 0 ? 1.0f : 0.0f;
    cam[3 + 13] = 0.795f;
    cam[3 + 14] = 3.4f;

    const uint32_t w = 24, hh = 18, spp = 2;
    std::vector<float> rgba(4 * w * hh);
    std::vector<double> rgba64(4 * w * hh);
    const int tracers[] = {0, 1, 2, 3, 4, 5, 6, 16};
    for (int tr : tracers) {
        orc_render(h, cam, tr, 5, 6, 7, 1, w, hh, spp, nullptr, w * hh, 0, 4, rgba.data(), nullptr);
        bool finite = true;
        for (float x : rgba)
            finite = finite && std::isfinite(x);
        expect(finite, "tracer " + std::to_string(tr) + ": non-finite film");
        orc_render(h, cam, tr, 5, 6, 7, 0, w, hh, 1, nullptr, w * hh, 1, 2, rgba64.data(), nullptr);
        orc_render_firefly(h, cam, tr, 5, 6, 7, w, hh, spp, 3, rgba.data());
    }

    // Ray queries of a few seeded rays through the box.
    std::vector<float> rays;
    for (int k = 0; k < 64; k++) {
        float a = 0.1f * k;
        rays.insert(rays.end(), {0.0f, 0.8f, 0.5f, std::cos(a), std::sin(0.7f * a), -0.5f});
    }
    const uint32_t nr = 64;
    std::vector<int32_t> geo(nr), tri(nr), prim(nr);
    std::vector<float> t(nr), b(2 * nr), vtx(3 * nr), nrm(3 * nr);
    orc_intersect(h, rays.data(), nr, 1, geo.data(), tri.data(), prim.data(), t.data(), b.data(), vtx.data(),
                  nrm.data(), nullptr);
    int hits = 0;
    for (int32_t g : geo)
        hits += g >= 0;
    expect(hits > 32, "too few hits from inside the box");
    orc_scene_free(h);

    // Compositors on a film with special values, and the math helpers.
    std::vector<float> film(4 * w * hh);
    for (size_t p = 0; p < film.size(); p++)
        film[p] = static_cast<float>(p % 17) * 0.37f - 1.0f;
    film[5] = NAN;
    film[9] = INFINITY;
    std::vector<uint8_t> px(4 * w * hh);
    float e = 0;
    uint32_t cnt = 0;
    for (int kind = 0; kind < 2; kind++)
        for (int autoe = 0; autoe < 2; autoe++) {
            orc_compose(film.data(), w, hh, kind, autoe, 0.7f, px.data(), &e, &cnt);
            orc_compose_libm(film.data(), w, hh, kind, autoe, 0.7f, px.data(), &e, &cnt);
        }
    std::vector<float> x = {0.0f, 1e-45f, 0.5f, 1.0f, 3.0f, -2.0f, NAN, INFINITY}, y(x.size());
    for (int fn = 0; fn < 3; fn++)
        orc_compose_math(fn, x.data(), static_cast<uint32_t>(x.size()), y.data());

    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("sanitize_check ok: OBJ reader (8 files, %zu malformed), oracle scene (%u nodes), %zu tracers, "
                "compositors\n",
                sizeof bad / sizeof bad[0], nn, sizeof tracers / sizeof tracers[0]);
    return 0;
}
