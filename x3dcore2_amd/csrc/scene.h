// Host-side scene: validation, BVH build in the reference's layout, device packing.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "scene_layout.h"

namespace x3d {

struct HostVec3 {
    float x, y, z;
};

struct HostMesh {
    std::vector<HostVec3> verts, normals;
    std::vector<uint32_t> tris;
    int32_t material = -1;
    int32_t light = -1;
};

struct HostLight {
    int32_t mesh; // -1: the sky (sky_light, light.cpp:87-163)
    float rad[3];
};

// Flattened node in the reference layout (pathspace.h:130-150), kept for dumps and tests.
struct FlatNode {
    float mn[3], mx[3];
    uint32_t num_prims, split_axis, next_child, prim_start;
};

struct PrimRef {
    uint32_t mesh, tri;
};

struct DeviceBuffers {
    void *nodes = nullptr, *nodes4 = nullptr, *tris = nullptr, *tri_attr = nullptr, *mats = nullptr,
         *lights = nullptr, *light_cum_power = nullptr, *light_tris = nullptr, *light_tri_cum = nullptr;
    uint64_t bytes = 0;
};

// Objects loaded by id, as objdb::push_updates hands them to the reference's actuators
// (obj.h:46-55, objdb.cpp:59-87): geometries, materials and lights in maps keyed by obj_id_t.
struct KeyedGeometry {
    HostMesh mesh;
    bool has_material = false;
    uint64_t material = 0; // obj id of the material (trimesh::material_id)
};

struct KeyedLight {
    uint64_t geometry; // obj id of the emitting geometry
    float rad[3];
    bool sky = false; // sky_light: no geometry
};

// trimesh::transform (geometry.cpp:76-87) with a column-major 4x4 matrix m[4 * col + row]:
// positions (M (x, y, z, 1)).cart(), normals (~M^-1 (n, 1)).cart(), in the float arithmetic of
// e8util::mat (tensor.h:310-560): the inverse is the adjugate (cofactor expansion along the first
// remaining column) scaled by 1/det. Returns "" or an error (matrix not invertible).
std::string transform_mesh(float const m[16], std::vector<HostVec3> &verts, std::vector<HostVec3> &normals);

class Scene {
  public:
    ~Scene();

    std::vector<DevMaterial> materials; // as added; fail-safe appended at commit
    std::vector<HostMesh> meshes;
    std::vector<HostLight> lights;
    bool committed = false;

    // Results of commit().
    std::vector<FlatNode> flat;
    std::vector<PrimRef> prims;
    uint32_t max_leaf_depth = 0;
    uint32_t num_interior = 0;
    uint32_t num_nodes4 = 0, max_stack4 = 0; // 4-wide traversal tree (see upload())
    double build_seconds = 0, upload_seconds = 0;
    DevScene dev{};
    DeviceBuffers buf;

    // Objects by id (x3d_scene_load_*): flattened in id order at every commit.
    bool keyed = false;
    std::map<uint64_t, KeyedGeometry> k_geos;
    std::map<uint64_t, DevMaterial> k_mats;
    std::map<uint64_t, KeyedLight> k_lights;
    std::vector<uint64_t> geo_ids; // obj id of each committed mesh (keyed scenes)
    std::vector<uint64_t> mat_ids; // obj id of each committed material (keyed scenes)

    // if_path_space::m_bound (pathspace.cpp:13-19): the union of every geometry ever loaded; it
    // grows with each load and is never reduced.
    bool space_empty = true;
    float space_lo[3] = {0, 0, 0}, space_hi[3] = {0, 0, 0};
    void grow_space(std::vector<HostVec3> const &verts);

    // Builds everything and uploads; returns an empty string or an error message. A keyed
    // scene may be committed again after edits: the previous device copy is released first.
    std::string commit();

  private:
    void build_bvh();
    std::string upload();
    std::string flatten_keyed();
    void release();
};

} // namespace x3d
