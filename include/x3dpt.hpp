/*
 * x3dpt.hpp — header-only C++ layer over the C ABI of libx3dpt.so (x3dpt.h): owning handles
 * that release themselves, std::vector buffers, and an exception (x3d::error, carrying the
 * status code and x3d_last_error()) instead of status codes. Nothing here adds behaviour; every
 * call forwards to the C entry point of the same name.
 */
#ifndef X3DPT_HPP
#define X3DPT_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "x3dpt.h"

namespace x3d {

class error : public std::runtime_error {
  public:
    explicit error(int code) : std::runtime_error(x3d_last_error()), code_(code) {}
    int code() const noexcept { return code_; }

  private:
    int code_;
};

inline void check(int rc) {
    if (rc != X3D_OK)
        throw error(rc);
}

// Move-only owner of a C handle with its destroy function.
template <class T, int (*Destroy)(T *)> class handle {
  public:
    handle() = default;
    explicit handle(T *p) : p_(p) {}
    handle(handle const &) = delete;
    handle &operator=(handle const &) = delete;
    handle(handle &&o) noexcept : p_(std::exchange(o.p_, nullptr)) {}
    handle &operator=(handle &&o) noexcept {
        if (this != &o) {
            reset();
            p_ = std::exchange(o.p_, nullptr);
        }
        return *this;
    }
    ~handle() { reset(); }
    T *get() const noexcept { return p_; }
    void reset() noexcept {
        if (p_)
            Destroy(p_);
        p_ = nullptr;
    }

  private:
    T *p_ = nullptr;
};

// The scene actuators (x3d_scene_*).
class scene {
  public:
    scene() {
        x3d_scene *s = nullptr;
        check(x3d_scene_create(&s));
        h_ = handle<x3d_scene, x3d_scene_destroy>(s);
    }
    x3d_scene *get() const noexcept { return h_.get(); }

    int32_t add_material(x3d_material_desc const &d) {
        int32_t id = -1;
        check(x3d_scene_add_material(get(), &d, &id));
        return id;
    }
    int32_t add_mesh(std::vector<float> const &verts, std::vector<float> const &normals,
                     std::vector<uint32_t> const &tris, int32_t material = -1) {
        int32_t id = -1;
        check(x3d_scene_add_mesh(get(), verts.data(), normals.data(), nullptr,
                                 static_cast<uint32_t>(verts.size() / 3), tris.data(),
                                 static_cast<uint32_t>(tris.size() / 3), material, &id));
        return id;
    }
    int32_t add_obj(std::string const &path, int32_t material = -1) {
        int32_t id = -1;
        check(x3d_scene_add_obj(get(), path.c_str(), material, &id));
        return id;
    }
    int32_t add_area_light(int32_t mesh, float const radiance[3]) {
        int32_t id = -1;
        check(x3d_scene_add_area_light(get(), mesh, radiance, &id));
        return id;
    }
    // Objects by id (the if_obj_actuator protocol).
    void load_geometry(uint64_t id, std::vector<float> const &verts, std::vector<float> const &normals,
                       std::vector<uint32_t> const &tris, bool has_material, uint64_t material,
                       float const *transform = nullptr) {
        check(x3d_scene_load_geometry(get(), id, verts.data(), normals.data(), nullptr,
                                      static_cast<uint32_t>(verts.size() / 3), tris.data(),
                                      static_cast<uint32_t>(tris.size() / 3), has_material ? 1 : 0, material,
                                      transform));
    }
    void load_material(uint64_t id, x3d_material_desc const &d) { check(x3d_scene_load_material(get(), id, &d)); }
    void load_light(uint64_t id, uint64_t geometry, float const radiance[3]) {
        check(x3d_scene_load_light(get(), id, geometry, radiance));
    }
    void unload(int32_t protocol, uint64_t id) { check(x3d_scene_unload(get(), protocol, id)); }
    void commit() { check(x3d_scene_commit(get())); }
    x3d_scene_info info() const {
        x3d_scene_info i{};
        check(x3d_scene_get_info(get(), &i));
        return i;
    }

  private:
    handle<x3d_scene, x3d_scene_destroy> h_;
};

// pt_image_renderer + pathtracer_factory (x3d_renderer_*, x3d_render*).
class renderer {
  public:
    explicit renderer(x3d_render_options const &o) {
        x3d_renderer *r = nullptr;
        check(x3d_renderer_create(&o, &r));
        h_ = handle<x3d_renderer, x3d_renderer_destroy>(r);
    }
    static x3d_render_options defaults(int32_t pt_type) {
        x3d_render_options o{};
        check(x3d_render_options_default(pt_type, &o));
        return o;
    }
    x3d_renderer *get() const noexcept { return h_.get(); }

    // The film: rgba float[4*w*h], pixel (x, y) at x + y*w.
    std::vector<float> render(scene const &s, x3d_camera const &cam, uint32_t w, uint32_t h, uint32_t spp,
                              bool firefly_filter = false, x3d_render_stats *stats = nullptr) {
        std::vector<float> film(4ull * w * h);
        check(x3d_render(get(), s.get(), &cam, w, h, spp, firefly_filter ? 1 : 0, film.data(), stats));
        return film;
    }
    void render_device(scene const &s, x3d_camera const &cam, uint32_t w, uint32_t h, uint32_t spp, float *d_rgba,
                       x3d_render_stats *stats = nullptr) {
        check(x3d_render_device(get(), s.get(), &cam, w, h, spp, d_rgba, stats));
    }
    x3d_autotune_report autotune(scene const &s, x3d_camera const &cam, uint32_t w, uint32_t h, uint32_t probe_spp) {
        x3d_autotune_report rep{};
        check(x3d_renderer_autotune(get(), s.get(), &cam, w, h, probe_spp, &rep));
        return rep;
    }

  private:
    handle<x3d_renderer, x3d_renderer_destroy> h_;
};

// clamp_compositor / aces_compositor (x3d_compositor_*).
class compositor {
  public:
    explicit compositor(int32_t kind = X3D_COMPOSITOR_ACES) {
        x3d_compositor *c = nullptr;
        check(x3d_compositor_create(kind, &c));
        h_ = handle<x3d_compositor, x3d_compositor_destroy>(c);
    }
    x3d_compositor *get() const noexcept { return h_.get(); }
    void enable_auto_exposure(bool on) { check(x3d_compositor_enable_auto_exposure(get(), on ? 1 : 0)); }
    void exposure(float e) { check(x3d_compositor_set_exposure(get(), e)); }
    // The 8-bit frame: r, g, b, a per pixel, pixel (x, y) at x + y*w.
    std::vector<uint8_t> commit(std::vector<float> const &film, uint32_t w, uint32_t h,
                                x3d_compose_stats *stats = nullptr) {
        std::vector<uint8_t> px(4ull * w * h);
        check(x3d_compositor_commit(get(), film.data(), w, h, px.data(), stats));
        return px;
    }

  private:
    handle<x3d_compositor, x3d_compositor_destroy> h_;
};

} // namespace x3d

#endif /* X3DPT_HPP */
