// Device-resident scene layout shared by the host packer (scene.cpp) and the HIP kernels.
//
// HBM layout (all arrays 16-byte aligned):
//   nodes     4 x float4 per interior BVH node (64 B): the two child boxes and child refs
//             f0 = (c0.min.x, c0.max.x, c0.min.y, c0.max.y)
//             f1 = (c1.min.x, c1.max.x, c1.min.y, c1.max.y)
//             f2 = (c0.min.z, c0.max.z, c1.min.z, c1.max.z)
//             f3 = (ref c0, ref c1, -, -) as uint32 bits; c0 = left child (flattened n + 1),
//             c1 = right child (flattened next_child) of the reference layout.
//   nodes4    kWideRows x float4 per wide node (128 B for 4 children): a tree over the
//             reference layout's leaves (inner nodes built for traversal cost, children
//             smallest surface area first): (min x)[W], (max x)[W], (min y)[W], (max y)[W],
//             (min z)[W], (max z)[W] of the children, their refs (kDone4 = empty slot, whose
//             box is [+inf, +inf] on every axis), padding.
//   tris      3 x float4 per primitive, in BVH primitive order: v0 (w = mesh id bits),
//             v1 (w = triangle index bits), v2 (w = 0).
//   tri_attr  3 x float4 per primitive: n0 (w = material id bits), n1 (w = light id bits or
//             -1), n2 (w = 0) — the shading normals read once per hit.
//   light_tris 6 x float4 per emitter triangle: p0, p1, p2, n0, n1, n2.
#pragma once

#include <stdint.h>

namespace x3d {

constexpr int32_t kMatOrenNayar = 0;
constexpr int32_t kMatCookTorr = 1;
constexpr int32_t kMatMixture = 2;
constexpr int32_t kMatFailSafe = 3;

struct DevMaterial { // 64 B
    int32_t kind;
    float albedo[3];
    float on_a, on_b; // Oren-Nayar A and B (material.cpp:116-118)
    float alpha2;     // cook_torr alpha^2 = 2 r^2 (material.cpp:379)
    float ior_re, ior_im;
    float ratio;
    int32_t sub0, sub1;
    float roughness, sigma;
    float sample_alpha2; // alpha^2 the GGX lobe is sampled with (and its pdf); alpha2 unless the
                         // evaluated parameters were changed with update_sampling off
    int32_t pad;
};
static_assert(sizeof(DevMaterial) == 64, "DevMaterial layout");

constexpr uint32_t kLightArea = 0; // area_light (light.cpp:17-85)
constexpr uint32_t kLightSky = 1;  // sky_light (light.cpp:87-163)
constexpr uint32_t kNoMesh = 0xFFFFFFFFu;

struct DevLight { // 64 B
    uint32_t tri_begin, tri_count;
    float area; // sum of triangle areas (geometry.cpp:105-118); 0 for the sky
    float prob; // |power| / total power (lightsources.cpp:73); 0 when the total is 0
    float rad[3];
    uint32_t mesh; // kNoMesh for the sky
    uint32_t kind; // kLightArea / kLightSky
    float sky_dia; // sky_light::m_dia, set at commit by set_scene_boundary (light.cpp:96-100)
    float sky_ref[3]; // sky_light::m_ref_p: centroid of the path-space bounds with z = 0
    uint32_t pad[3];
};
static_assert(sizeof(DevLight) == 64, "DevLight layout");

// Child reference encoding: interior node index, or a leaf:
// bit 31 = leaf, bits 28..29 = primitive count - 1 (1..4 primitives), bits 0..27 = first prim.
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmptyRef = 0xFFFFFFFFu; // scene without triangles
constexpr uint32_t kMaxPrims = 1u << 28;

inline constexpr uint32_t leaf_ref(uint32_t start, uint32_t count) {
    return kLeafBit | ((count - 1u) << 28) | start;
}

constexpr uint32_t kDone4 = 0xFFFFFFFFu; // empty child slot of a wide node
// Children per node of the wide traversal tree, and float4 rows per node: for each of the six
// planes (min x, max x, min y, max y, min z, max z) W/4 rows of four children, then W/4 rows of
// child refs, padded to a power of two (128 B for W = 4).
constexpr uint32_t kTreeWidth = 4;
constexpr uint32_t kWideGroups = kTreeWidth / 4;
constexpr uint32_t kWideRows = kTreeWidth == 4 ? 8 : 16;
static_assert(kTreeWidth == 4 || kTreeWidth == 8, "wide tree of 4 or 8 children");
// Wide nodes numbered first in breadth-first order (the top of the tree), which the wavefront
// traversal kernels copy into LDS per workgroup; 0 = none. Staging measured slower on the
// spheres workload (16 nodes: extend +12%, 32 nodes: +24%): the extra LDS per workgroup costs
// resident waves, the top nodes hit in L1 / L2 anyway, and lanes at different depths execute
// both load paths. The mechanism stays for scenes where it may pay.
constexpr uint32_t kTopNodes = 0;

struct DevScene {
    const void *nodes;      // float4[4 * num_interior]
    const void *nodes4;     // float4[kWideRows * wide nodes]: the wide traversal tree
    const void *tris;       // float4[3 * num_prims]
    const void *tri_attr;   // float4[3 * num_prims]
    const DevMaterial *mats;
    const DevLight *lights;
    const float *light_cum_power; // cumulative |power| per light
    const void *light_tris;       // float4[6 * total emitter triangles]
    const float *light_tri_cum;   // cumulative area per emitter triangle (per light)
    uint32_t root_ref;
    uint32_t root4;         // root of the wide tree (a leaf ref when the root is a leaf)
    uint32_t num_nodes4;    // wide nodes
    uint32_t num_lights;
    uint32_t num_mats;
    uint32_t fail_safe_mat; // index of the fail-safe Lambertian 0.8
    float total_power;
    uint32_t max_stack;     // stack entries a traversal of either tree can need
    uint32_t num_prims;
    uint32_t num_interior;
    float box_lo[3];  // scene bounds (root box): lower corner
    float box_inv[3]; // 1 / extent per axis (0 for a flat axis)
    uint32_t far_coords; // every ray traverses the binary tree: some box coordinate exceeds 1e30
                         // in magnitude (slab_may_nan), or the wide tree spans 4 GiB or more (its
                         // rows are addressed by 32-bit byte offsets from `nodes4`)
    // if_path_space::aabb (pathspace.cpp:13-18): union of the vertex bounds of every mesh
    // (trimesh::update_aabb, geometry.cpp:97-103); lower corner and extent max - min
    float space_lo[3];
    float space_range[3];
};

} // namespace x3d
