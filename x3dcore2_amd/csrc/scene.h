// Host-side scene: validation, BVH build in the reference's layout, device packing.
#pragma once

#include <cstdint>
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
    int32_t mesh;
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

    // Builds everything and uploads; returns an empty string or an error message.
    std::string commit();

  private:
    void build_bvh();
    std::string upload();
};

} // namespace x3d
