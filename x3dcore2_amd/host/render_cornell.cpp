// C++ host program over the C ABI (through include/x3dpt.hpp): the reference's Cornell box as
// tst_renderer builds it (tst_renderer.cpp:37-115) — its eight OBJ parts, Oren-Nayar walls
// (sigma 0.078), a cook_torr sphere (r 0.2, eta 2.93 + 3i), the ceiling light of radiance
// (0.911, 0.660, 0.345) x 15, the pinhole camera at (0, 0.795, 3.4) — rendered into the linear
// film and composed to the 8-bit frame with the ACES compositor at the pipeline's defaults (auto
// exposure off, exposure 1; pipeline.cpp:116-117), as pt_render_pipeline::render_frame does
// (pipeline.cpp:73-101).
//
//   render_cornell OBJ_DIR OUT_PREFIX WIDTH HEIGHT SPP TRACER [SEED [FORM]]
//     TRACER: lt1 | unidirect | mis | bdpt | lt2
//     FORM: plain (default) | wavefront (lt1, unidirect, mis; the same film)
//     writes OUT_PREFIX.film (float32 rgba), OUT_PREFIX.rgba (uint8 rgba), OUT_PREFIX.ppm
//   render_cornell OBJ_DIR --host-only
//     builds the scene on the host only and reports the device error of the commit (for
//     machines without a GPU)

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/x3dpt.hpp"

namespace {

x3d_material_desc oren_nayar(float r, float g, float b, float sigma) {
    x3d_material_desc d{};
    d.kind = X3D_MAT_OREN_NAYAR;
    d.albedo[0] = r;
    d.albedo[1] = g;
    d.albedo[2] = b;
    d.sigma = sigma;
    return d;
}

x3d_material_desc cook_torr(float r, float g, float b, float roughness, float eta, float k) {
    x3d_material_desc d{};
    d.kind = X3D_MAT_COOK_TORR;
    d.albedo[0] = r;
    d.albedo[1] = g;
    d.albedo[2] = b;
    d.roughness = roughness;
    d.ior_re = eta;
    d.ior_im = k;
    return d;
}

// Builds the box: returns the scene, uncommitted.
x3d::scene cornell(std::string const &dir) {
    x3d::scene s;
    int32_t white = s.add_material(oren_nayar(0.725f, 0.710f, 0.680f, 0.078f));
    int32_t red = s.add_material(oren_nayar(0.630f, 0.065f, 0.050f, 0.078f));
    int32_t green = s.add_material(oren_nayar(0.140f, 0.450f, 0.091f, 0.078f));
    int32_t glossy = s.add_material(cook_torr(0.95f, 0.95f, 0.95f, 0.2f, 2.93f, 3.0f));
    int32_t black = s.add_material(oren_nayar(0.0f, 0.0f, 0.0f, 0.078f));
    struct Part {
        const char *name;
        int32_t mat;
    } parts[] = {{"left_wall", red},        {"right_wall", green}, {"back_wall", white},
                 {"ceiling", white},        {"floor", white},      {"left_sphere", glossy},
                 {"right_sphere", white},   {"light", black}};
    int32_t light_mesh = -1;
    for (Part const &p : parts)
        light_mesh = s.add_obj(dir + "/" + p.name + ".obj", p.mat);
    float rad[3] = {0.911f * 15.0f, 0.660f * 15.0f, 0.345f * 15.0f};
    s.add_area_light(light_mesh, rad);
    return s;
}

x3d_camera camera() {
    x3d_camera c{};
    c.sensor_size = 0.032f;
    c.focal_length = 0.035f;
    c.aspect = 4.0f / 3.0f;
    for (int k = 0; k < 16; k++)
        c.transform[k] = k % 5 == 0 ? 1.0f : 0.0f;
    c.transform[13] = 0.795f; // column-major: translation in the last column
    c.transform[14] = 3.4f;
    return c;
}

template <class T> void write_raw(std::string const &path, std::vector<T> const &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

} // namespace

int main(int argc, char **argv) {
    if (argc == 3 && std::strcmp(argv[2], "--host-only") == 0) {
        try {
            x3d::scene s = cornell(argv[1]);
            try {
                s.commit();
                std::printf("committed (a device is present)\n");
            } catch (x3d::error const &e) {
                std::printf("host scene built; commit: error %d: %s\n", e.code(), e.what());
                return e.code() == X3D_ERR_DEVICE ? 0 : 1;
            }
        } catch (x3d::error const &e) {
            std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
            return 1;
        }
        return 0;
    }
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s OBJ_DIR OUT_PREFIX WIDTH HEIGHT SPP lt1|unidirect|mis|bdpt|lt2 [SEED [FORM]]\n",
                     argv[0]);
        return 2;
    }
    const std::string dir = argv[1], out = argv[2], tracer = argv[6];
    const uint32_t w = static_cast<uint32_t>(std::atoi(argv[3])), h = static_cast<uint32_t>(std::atoi(argv[4])),
                   spp = static_cast<uint32_t>(std::atoi(argv[5]));
    const uint64_t seed = argc > 7 ? std::strtoull(argv[7], nullptr, 10) : 1;
    const std::string form = argc > 8 ? argv[8] : "plain";
    if (form != "plain" && form != "wavefront") {
        std::fprintf(stderr, "unknown form %s\n", form.c_str());
        return 2;
    }
    int32_t pt = tracer == "lt1"         ? X3D_PT_UNIDIRECT_LT1
                 : tracer == "unidirect" ? X3D_PT_UNIDIRECT
                 : tracer == "mis" ? X3D_PT_UNIDIRECT_MIS
                 : tracer == "bdpt" ? X3D_PT_BIDIRECT_MIS
                 : tracer == "lt2" ? X3D_PT_BIDIRECT_LT2
                                   : -1;
    if (pt < 0) {
        std::fprintf(stderr, "unknown tracer %s\n", tracer.c_str());
        return 2;
    }
    try {
        x3d::scene s = cornell(dir);
        s.commit();
        x3d_render_options o = x3d::renderer::defaults(pt);
        o.seed = seed;
        o.form = form == "wavefront" ? X3D_FORM_WAVEFRONT : X3D_FORM_PLAIN;
        x3d::renderer r(o);
        x3d_render_stats st{};
        std::vector<float> film = r.render(s, camera(), w, h, spp, false, &st);
        x3d::compositor com(X3D_COMPOSITOR_ACES);
        com.enable_auto_exposure(false);
        com.exposure(1.0f);
        x3d_compose_stats cs{};
        std::vector<uint8_t> px = com.commit(film, w, h, &cs);
        write_raw(out + ".film", film);
        write_raw(out + ".rgba", px);
        std::ofstream ppm(out + ".ppm", std::ios::binary);
        ppm << "P6\n" << w << " " << h << "\n255\n";
        for (size_t p = 0; p < static_cast<size_t>(w) * h; p++)
            ppm.write(reinterpret_cast<const char *>(&px[4 * p]), 3);
        std::printf("%ux%u x %u spp (%s, %s): %.2f ms render, exposure %.5f, compose %.3f ms\n", w, h, spp,
                    tracer.c_str(), form.c_str(), st.total_ms, cs.exposure, cs.exposure_ms + cs.compose_ms);
    } catch (x3d::error const &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
