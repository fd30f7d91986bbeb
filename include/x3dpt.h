/*
 * x3dpt.h — C ABI of the MI355X (gfx950) path-tracing library, libx3dpt.so.
 *
 * Drop-in boundary for the per-sample path-tracing loop of DaviesX/x3dcore2 ("e8yescg"):
 * scene + camera + spp in, linear-RGB film out. Plain pointers and sizes, opaque handles,
 * int status codes (X3D_OK = 0) with a thread-local message from x3d_last_error(); no C++
 * exception crosses the boundary. All host buffers are caller-owned. One call in flight per
 * handle; distinct handles are independent.
 *
 * Each entry point names the reference interface it replaces (file:line under corelib/src or
 * corelib/test of the reference).
 */
#ifndef X3DPT_H
#define X3DPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define X3D_OK 0
#define X3D_ERR_INVALID_ARGUMENT 1 /* null/empty/out-of-range input */
#define X3D_ERR_STATE 2            /* e.g. scene not committed, edit after commit */
#define X3D_ERR_DEVICE 3           /* HIP runtime failure or no GPU */
#define X3D_ERR_OUT_OF_MEMORY 4
#define X3D_ERR_UNSUPPORTED 5

/* Last error message of the calling thread ("" when none). */
const char *x3d_last_error(void);

/* Library version string. */
const char *x3d_version(void);

/* ---- scene (replaces the if_obj_actuator back-ends bvh_path_space_layout,
 *      default_material_container and basic_light_sources: pathspace.h:40-58, 76-164,
 *      materialcontainer.h:14-53, lightsources.h:18-61) ---------------------------------- */
typedef struct x3d_scene x3d_scene;

/* Material kinds. OREN_NAYAR with sigma = 0 is the Lambertian of oren_nayar
 * (material.cpp:108-169); COOK_TORR is cook_torr (GGX D, Smith G, complex-IOR Fresnel,
 * alpha^2 = 2 r^2; material.cpp:171-252); MIXTURE is mat_mixture (material.cpp:78-106). */
#define X3D_MAT_OREN_NAYAR 0
#define X3D_MAT_COOK_TORR 1
#define X3D_MAT_MIXTURE 2

typedef struct x3d_material_desc {
    int32_t kind;
    float albedo[3];
    float sigma;     /* OREN_NAYAR roughness (0 = Lambertian) */
    float roughness; /* COOK_TORR r */
    float ior_re, ior_im;
    int32_t sub0, sub1; /* MIXTURE operands (material ids already added) */
    float ratio;        /* MIXTURE probability of sub0 */
} x3d_material_desc;

int x3d_scene_create(x3d_scene **out);
int x3d_scene_destroy(x3d_scene *scene);

/* Adds a material; *out_id is its id (0, 1, 2, ... in call order). */
int x3d_scene_add_material(x3d_scene *scene, const x3d_material_desc *desc, int32_t *out_id);

/* Adds a world-space triangle mesh (replaces if_path_space::load of a trimesh,
 * pathspace.cpp:15-19 / geometry.h:81-116). verts, normals: float[3*num_verts] (per-vertex
 * shading normals); uvs: optional float[2*num_verts] (may be NULL); tris: uint32[3*num_tris].
 * material_id < 0 selects the fail-safe Lambertian 0.8 (materialcontainer.cpp:7-17).
 * Meshes are numbered 0, 1, 2, ... in call order; that number is the hit's `geo`. */
int x3d_scene_add_mesh(x3d_scene *scene, const float *verts, const float *normals, const float *uvs,
                       uint32_t num_verts, const uint32_t *tris, uint32_t num_tris,
                       int32_t material_id, int32_t *out_mesh_id);

/* ---- Wavefront OBJ meshes (replaces e8util::wavefront_obj::load_geometry,
 * resource.cpp:137-255). Lines "v x y z", "vn x y z", "vt u v" and triangle faces
 * "f a/b/c a/b/c a/b/c" (1-based position / texcoord / normal indices; texcoord and normal may
 * be empty) are read with the reference's pattern rules; other lines are skipped. Vertices keep
 * their file order and take the normal and texcoord of the last face corner naming them. A face
 * with more than three corners, a corner without exactly three '/'-separated fields, a missing
 * or out-of-range index, a field that is not a number, or a file without vertices returns
 * X3D_ERR_INVALID_ARGUMENT with the line in x3d_last_error(). ---- */
typedef struct x3d_obj_mesh {
    uint32_t num_verts, num_tris;
    float *verts;    /* float[3 * num_verts] */
    float *normals;  /* float[3 * num_verts]; zero where no face corner gives one */
    float *uvs;      /* float[2 * num_verts]; zero where no face corner gives one */
    uint32_t *tris;  /* uint32[3 * num_tris], 0-based */
} x3d_obj_mesh;

int x3d_obj_load(const char *path, x3d_obj_mesh *mesh); /* arrays owned by the library */
int x3d_obj_free(x3d_obj_mesh *mesh);

/* x3d_obj_load + x3d_scene_add_mesh: the reference's load_geometry() + path_space::load() with
 * the identity transform, as in tst_renderer.cpp:58-102. */
int x3d_scene_add_obj(x3d_scene *scene, const char *path, int32_t material_id, int32_t *out_mesh_id);

/* Makes mesh `mesh_id` an area light of the given radiance (replaces area_light,
 * light.cpp:17-85, loaded into basic_light_sources, lightsources.cpp:11-36). */
int x3d_scene_add_area_light(x3d_scene *scene, int32_t mesh_id, const float radiance[3],
                             int32_t *out_light_id);

/* Adds the sky light (replaces sky_light, light.cpp:87-163, which the reference's glTF loader
 * adds as its virtual light, resource.cpp:518-523). It has no geometry: rays never hit it, and
 * it is reached only by light sampling. At x3d_scene_commit it is placed by
 * sky_light::set_scene_boundary (light.cpp:96-100) with the path space's bounds (the union of
 * every mesh's vertex bounds): diameter = 2 * (|max - min| / 2), reference point = the bounds'
 * centroid with z = 0 (the reference never calls set_scene_boundary, leaving both unset). A
 * light sample picks u cosine-distributed about +z, the point (u + ref) * dia with normal -u and
 * density (z . u) / pi; it sends L max(cos_light, 0) max(cos_target, 0) (no distance
 * fall-off). Its power, for the light choice, is pi * dia^2 / 2 per channel. In
 * X3D_PT_UNIDIRECT_MIS its light samples carry MIS weight 1 (no BSDF sample can reach it). */
int x3d_scene_add_sky_light(x3d_scene *scene, const float radiance[3], int32_t *out_light_id);

/* ---- objects by id (replace the three if_obj_actuator back-ends — if_path_space,
 * if_material_container, if_light_sources — as objdb::push_updates drives them with
 * load(obj, trans) / unload(obj) / commit(): obj.h:46-55, objdb.cpp:59-87,
 * pathspace.cpp:13-26, materialcontainer.cpp:20-33, lightsources.cpp:11-24).
 * A scene takes either these calls or the x3d_scene_add_* calls, not both. Objects are keyed by
 * the caller's obj_id_t; as in the reference's maps, loading an id that is already loaded of
 * the same protocol changes nothing (objdb unloads a dirty object before it reloads it), and
 * x3d_scene_commit orders geometries, materials and lights by id (the hit's `geo` is the
 * geometry's rank in id order; x3d_scene_geometry_obj_id maps it back). Edits are staged: renders
 * use the last commit, and a keyed scene may be committed again after edits (a failed commit
 * leaves it uncommitted). ---- */
#define X3D_OBJ_GEOMETRY 1 /* obj_protocol_geometry */
#define X3D_OBJ_LIGHT 3    /* obj_protocol_light */
#define X3D_OBJ_MATERIAL 4 /* obj_protocol_material */

/* if_path_space::load of a trimesh. transform: column-major 4x4 (NULL = identity), applied with
 * trimesh::transform's float arithmetic (x3d_transform_mesh). The geometry's material is the
 * material loaded under material_obj_id when has_material != 0 and that material is loaded at
 * commit, else the fail-safe Lambertian 0.8 (if_material_container::find,
 * materialcontainer.cpp:7-17). As in the reference, the path-space bounds that
 * X3D_PT_POSITION uses grow with every geometry load and never shrink. */
int x3d_scene_load_geometry(x3d_scene *scene, uint64_t obj_id, const float *verts, const float *normals,
                            const float *uvs, uint32_t num_verts, const uint32_t *tris, uint32_t num_tris,
                            int32_t has_material, uint64_t material_obj_id, const float *transform);
/* if_material_container::load (OREN_NAYAR, COOK_TORR; mixtures are added by index). */
int x3d_scene_load_material(x3d_scene *scene, uint64_t obj_id, const x3d_material_desc *desc);
/* if_light_sources::load of an area light on the geometry loaded under geometry_obj_id (which
 * must be loaded at commit). The light follows its geometry's transform (the reference's
 * area_light::transform keeps the untransformed geometry, light.cpp:83-85). */
int x3d_scene_load_light(x3d_scene *scene, uint64_t obj_id, uint64_t geometry_obj_id, const float radiance[3]);
/* if_light_sources::load of a sky light (see x3d_scene_add_sky_light; light.cpp:87-163),
 * placed at each commit with the path space's bounds at that time. */
int x3d_scene_load_sky_light(x3d_scene *scene, uint64_t obj_id, const float radiance[3]);
/* unload(obj) of the given protocol (X3D_OBJ_*); unknown ids are ignored. */
int x3d_scene_unload(x3d_scene *scene, int32_t protocol, uint64_t obj_id);
/* obj id of committed mesh `geo` (for add_* scenes: geo itself). */
int x3d_scene_geometry_obj_id(const x3d_scene *scene, int32_t geo, uint64_t *obj_id);

/* trimesh::transform (geometry.cpp:76-87) on the host: positions (M (x, y, z, 1)).cart(),
 * normals (~M^-1 (n, 1)).cart(), M^-1 = adjugate / det in e8util::mat's float arithmetic
 * (tensor.h:426-535). X3D_ERR_INVALID_ARGUMENT when |det| <= 1e-5 (the reference throws). */
int x3d_transform_mesh(const float transform[16], uint32_t num_verts, const float *verts, const float *normals,
                       float *out_verts, float *out_normals);

/* Builds the SAH BVH on the host with the reference's layout (bvh_path_space_layout::commit,
 * pathspace.cpp:292-328), the light power CDF (lightsources.cpp:45-52), and uploads the
 * scene to device memory. Beside the reference layout it builds the 4-wide traversal tree most
 * rays use: the reference's leaves under inner nodes of its own (a 3-axis SAH over the leaf
 * boxes); which leaves a ray reaches, and so every hit, does not depend on the inner nodes.
 * Scenes built with x3d_scene_add_* are immutable afterwards. */
int x3d_scene_commit(x3d_scene *scene);

/* Material parameters of a committed scene (no reference counterpart; for inverse rendering):
 * materials [first, first+count) get albedo[3*k..] (OREN_NAYAR, COOK_TORR) and roughness[k]
 * (COOK_TORR, > 0) when those arrays are non-NULL. update_sampling != 0 also moves the density
 * the GGX lobe is sampled with to the new roughness; with 0 the sampling density stays, so that
 * renders are smooth in the evaluated parameters for fixed seeds (the derivative that
 * x3d_render_backward returns). Mixtures have no parameters of their own (their entries are
 * skipped). The device table is updated synchronously before the call returns. */
int x3d_scene_set_material_params(x3d_scene *scene, uint32_t first, uint32_t count, const float *albedo,
                                  const float *roughness, int32_t update_sampling);

typedef struct x3d_scene_info {
    uint32_t num_meshes, num_tris, num_nodes, num_interior, num_lights, num_materials;
    uint32_t max_leaf_depth;   /* deepest leaf (root = depth 1) */
    uint64_t device_bytes;     /* scene bytes resident in HBM */
    double build_seconds;      /* host BVH build time */
    double upload_seconds;     /* host-to-device copy time */
} x3d_scene_info;
int x3d_scene_get_info(const x3d_scene *scene, x3d_scene_info *info);

/* Dumps the flattened BVH in the reference layout (flattened_node, pathspace.h:130-150):
 * bounds float[6*num_nodes] (min xyz, max xyz); meta uint32[4*num_nodes] (num_prims,
 * split_axis, next_child, prim_start); prims int32[2*num_tris] (mesh, triangle). */
int x3d_scene_dump_bvh(const x3d_scene *scene, float *bounds, uint32_t *meta, int32_t *prims);

/* ---- ray queries (replace bvh_path_space_layout::intersect / has_intersect,
 *      pathspace.cpp:330-456) -------------------------------------------------------------- */
typedef struct x3d_hit {
    int32_t geo;    /* mesh id, -1 on miss */
    int32_t tri;    /* triangle index within the mesh, -1 on miss */
    int32_t prim;   /* index into the BVH primitive array, -1 on miss */
    float t;        /* ray parameter of the hit (INFINITY on miss) */
    float b1, b2;   /* barycentrics of vertex 1 and 2 (tensor.h:688-725) */
    float vertex[3];
    float normal[3]; /* interpolated, normalized shading normal (pathspace.cpp:385-409) */
} x3d_hit;

/* Traversal modes. REFERENCE visits every box overlapping [1e-4, 1000] in the reference's
 * LIFO order (pathspace.cpp:343-383). CULLED (default) visits the nearer child first and skips
 * boxes that start beyond the best hit; ties keep the reference's winner. Both return the same
 * hits as the reference's intersector. */
#define X3D_TRAVERSAL_CULLED 0
#define X3D_TRAVERSAL_REFERENCE 1

/* Closest hits over t in [1e-4, 1000]. rays: float[6*n] (origin, direction), host memory. */
int x3d_intersect(const x3d_scene *scene, const float *rays, uint32_t n, int32_t traversal,
                  x3d_hit *hits);

/* Any hit with t in [tmin[k], tmax[k]] (has_intersect). occluded: uint8[n] (1 = blocked). */
int x3d_occluded(const x3d_scene *scene, const float *rays, const float *tmin, const float *tmax,
                 uint32_t n, uint8_t *occluded);

/* Device-pointer variants for callers that keep rays in HBM (no host copies): d_rays, d_hits
 * etc. are device pointers; the call is asynchronous on `stream` (a hipStream_t, or NULL for
 * the default stream). */
int x3d_intersect_device(const x3d_scene *scene, const float *d_rays, uint32_t n, int32_t traversal,
                         x3d_hit *d_hits, void *stream);
int x3d_occluded_device(const x3d_scene *scene, const float *d_rays, const float *d_tmin,
                        const float *d_tmax, uint32_t n, uint8_t *d_occluded, void *stream);

/* ---- BSDF batches (replace if_material::eval / sample, material.h:75-89; pdf is new:
 *      the solid-angle density sample() returns, as a function of (o, i)) ----------------- */
/* f: float[3*n]. normals, o, i: float[3*n]. */
int x3d_bsdf_eval(const x3d_scene *scene, int32_t material_id, uint32_t n, const float *normals,
                  const float *o, const float *i, float *f);
int x3d_bsdf_pdf(const x3d_scene *scene, int32_t material_id, uint32_t n, const float *normals,
                 const float *o, const float *i, float *pdf);
/* Sample k draws from the PCG stream (seed, pixel = k, sample = 0). reference_mixture = 1 uses
 * mat_mixture's density (chosen lobe's density times its probability, material.cpp:94-106);
 * 0 the full mixture density. i: float[3*n], dens: float[n]. */
int x3d_bsdf_sample(const x3d_scene *scene, int32_t material_id, uint64_t seed, uint32_t n,
                    const float *normals, const float *o, int32_t reference_mixture, float *i,
                    float *dens);

/* ---- camera (replaces pinhole_camera, camera.cpp:23-71) --------------------------------- */
typedef struct x3d_camera {
    float sensor_size, focal_length, aspect;
    float transform[16]; /* column-major 4x4; rotation = upper-left 3x3, position = column 3 */
} x3d_camera;

/* pinhole_camera::sample for every pixel (deterministic, no jitter). rays: float[6*w*h]. */
int x3d_camera_rays(const x3d_camera *camera, uint32_t width, uint32_t height, float *rays);

/* ---- renderer (replaces pt_image_renderer(factory, threads) + render(...),
 *      renderer.h:30, 57-59, and pathtracer_factory, pathtracerfact.h:12-21) -------------- */
typedef struct x3d_renderer x3d_renderer;

/* Tracer types: pathtracer_factory::pt_type values (pathtracerfact.h:12) ... */
#define X3D_PT_NORMAL 0         /* normal_tracer (pathtracer.cpp:492-506) */
#define X3D_PT_POSITION 1       /* position_tracer (pathtracer.cpp:474-490) */
#define X3D_PT_DIRECT 2         /* direct_path_tracer (pathtracer.cpp:508-526) */
#define X3D_PT_UNIDIRECT 3     /* unidirect_path_tracer (pathtracer.cpp:528-585) */
#define X3D_PT_UNIDIRECT_LT1 4 /* unidirect_lt1_path_tracer (pathtracer.cpp:587-653) */
#define X3D_PT_BIDIRECT_LT2 5   /* bidirect_lt2_path_tracer (pathtracer.cpp:655-752) */
#define X3D_PT_BIDIRECT_MIS 6  /* bidirect_mis_path_tracer (pathtracer.cpp:754-803) */
/* ... plus the fixed-depth unidirectional tracer with light + BSDF sampling combined by the
 * power heuristic (no reference counterpart). */
#define X3D_PT_UNIDIRECT_MIS 16

/* Execution forms of the same estimator; every form gives the bit-identical film.
 * x3d_render_options_default picks X3D_FORM_AUTO for UNIDIRECT_MIS, UNIDIRECT_LT1, UNIDIRECT
 * and BIDIRECT_MIS, the plain form for the other tracers. */
#define X3D_FORM_PLAIN 0     /* persistent kernels over the (pixel, sample) items of a generation,
                                each lane taking the next item when its path ends, then an in-order
                                per-pixel sum (the drop-in) */
#define X3D_FORM_WAVEFRONT 1 /* ray queues per bounce, bounce loop in a hipGraph; runs
                                UNIDIRECT_MIS, UNIDIRECT_LT1 and UNIDIRECT (the last two unsorted);
                                BIDIRECT_MIS: subpath vertices in structure-of-arrays rows, walk
                                rays per bounce, then every connection ray in one any-hit pass
                                (unsorted, no graph; generations of whole samples sized by
                                queue_capacity and a 24 GiB working-memory budget, about 2.4 KB
                                per (pixel, sample) item at 6-vertex subpaths) */
#define X3D_FORM_AUTO 2      /* per render: the plain form below a per-tracer size that grows as
                                the scene's triangle count falls, the wavefront form from it on
                                (w*h*spp samples; from 2^18 triangles on: UNIDIRECT_MIS 2^20,
                                UNIDIRECT_LT1 2^22, UNIDIRECT 2^23, BIDIRECT_MIS 2^21; for a few
                                dozen triangles: 2^24, 2^24, 2^24, 2^22 — where the two forms
                                measured equal); the plain form for the other tracers; batches
                                run the wavefront form */

typedef struct x3d_render_options {
    int32_t pt_type;       /* X3D_PT_* */
    int32_t form;          /* X3D_FORM_* */
    uint32_t max_depth;    /* UNIDIRECT_MIS: bounces after the first hit */
    uint32_t max_path_len; /* BIDIRECT_MIS: vertices per subpath (reference: 8) */
    uint64_t seed;         /* PCG seed; sample s of pixel p uses stream (seed, p, s) */
    uint32_t workgroup_size; /* WAVEFRONT: traversal workgroup size (64/128/256), 0 = default (64) */
    uint32_t queue_capacity; /* WAVEFRONT: paths in flight per generation, 0 = default = max = 64M
                              (about 270 B per path; a render allocates for min(capacity,
                              w*h*spp));
                              PLAIN (every tracer) and WAVEFRONT BIDIRECT_MIS: (pixel, sample)
                              items per generation, whole samples of every pixel (default 16M;
                              the sample buffer takes 16 B per item, 256 MiB at the default,
                              min(capacity, w*h*spp) items) */
    int32_t sort_key;        /* WAVEFRONT: X3D_SORT_* */
    uint32_t stack_depth;    /* WAVEFRONT: LDS traversal-stack entries per lane (8/16), 0 = 16;
                              16 gives the any-hit kernels 16 entries and the closest-hit kernel
                              12 (deeper entries spill to device memory either way) */
    int32_t profile;         /* WAVEFRONT: 1 = time traversal kernels with events (no graph) */
    uint32_t batch_split;    /* batches: jobs of at least this many samples (w*h*spp) run in the
                              wavefront form after the shared launch of the others; 0 = default
                              (4M), 0xFFFFFFFF = every job in the shared launch */
    int32_t ray_order;       /* WAVEFRONT: X3D_RAY_ORDER_* */
    uint32_t reserved;
} x3d_render_options;

/* Wavefront ray order before each traversal kernel: the order the rays were queued in (default),
 * or sorted by direction octant and origin cell (a counting sort per bounce; the traversal
 * kernels fetch rays dynamically, which already keeps their lanes busy, so the sort costs more
 * than it saves on the measured workloads). The film does not depend on it. */
#define X3D_RAY_ORDER_DEFAULT 0
#define X3D_RAY_ORDER_SORTED 1
#define X3D_RAY_ORDER_APPEND 2

/* Wavefront sort keys: keep append order (default) or group active paths by material kind. */
#define X3D_SORT_DEFAULT 0
#define X3D_SORT_MATERIAL 1
#define X3D_SORT_NONE 2

/* Fills *opts with the defaults for pt_type. */
int x3d_render_options_default(int32_t pt_type, x3d_render_options *opts);

int x3d_renderer_create(const x3d_render_options *opts, x3d_renderer **out);
int x3d_renderer_destroy(x3d_renderer *r);

typedef struct x3d_render_stats {
    uint64_t num_samples;       /* w*h*spp */
    uint64_t closest_queries;   /* closest-hit rays traced (including the first-hit pass) */
    uint64_t anyhit_queries;    /* shadow rays traced */
    uint64_t closest_nodes;     /* interior nodes visited by closest-hit queries */
    uint64_t closest_tests;     /* ray-triangle tests of closest-hit queries */
    uint64_t anyhit_nodes;      /* interior nodes visited by any-hit queries */
    uint64_t anyhit_tests;      /* ray-triangle tests of any-hit queries */
    uint64_t first_hit_nodes;   /* ... of which in the first-hit pass (closest_nodes) */
    uint64_t first_hit_tests;   /* ... of which in the first-hit pass (closest_tests) */
    double kernel_ms;           /* device time of all render kernels (HIP events) */
    double first_hit_ms;        /* device time of the first-hit pass */
    double trace_ms;            /* device time of the sample kernels (bounce loop) */
    double total_ms;            /* wall time of the call including copies */
    double extend_ms;           /* WAVEFRONT + profile: closest-hit traversal kernels, summed */
    double shadow_ms;           /* WAVEFRONT + profile: any-hit traversal kernels, summed */
    uint32_t trace_launches;    /* WAVEFRONT + profile: launches of each traversal kernel */
    int32_t form;               /* the form that rendered: X3D_FORM_PLAIN or X3D_FORM_WAVEFRONT */
} x3d_render_stats;

/* Autotuner (no reference counterpart): renders `probe_spp` samples per pixel of the given
 * workload with candidate wavefront configurations, keeps the fastest in the renderer's options
 * and reports it. Stage 1 tries traversal workgroup size {64,128,256} x LDS stack depth {8,16} x
 * sort key {material, none} at the default capacity; stage 2 tries queue capacities
 * {2M, 4M, 8M, 16M, 32M} paths for the stage-1 winner, skipping any above width*height*probe_spp
 * (a probe that small cannot tell them apart). Each candidate is timed as the best of two renders
 * after one warm-up render; the default configuration is kept unless the winner is at least 2%
 * faster. The film does not depend on the configuration. */
typedef struct x3d_autotune_report {
    uint32_t workgroup_size, stack_depth, queue_capacity;
    int32_t sort_key;
    double best_ms;      /* trace time of the probe render with the chosen configuration */
    double default_ms;   /* ... with the default configuration */
    uint32_t candidates; /* configurations tried */
    uint32_t reserved;
} x3d_autotune_report;

int x3d_renderer_autotune(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera,
                          uint32_t width, uint32_t height, uint32_t probe_spp,
                          x3d_autotune_report *report);

/* Renders the film: rgba float[4*w*h], pixel (x, y) at index x + y*w, row 0 = top; rgb is the
 * mean radiance over spp samples, alpha = 1/spp (pt_image_renderer::render with one task: renderer.cpp:140-159).
 * firefly_filter != 0 applies sampling_task::run's filter (renderer.cpp:46-83) to every sample
 * image before the per-pixel sum: an interior pixel whose squared norm exceeds each of its 8
 * neighbours' by more than 1.5 is capped, per component, at their mean squared norm + 1.5. The
 * filter acts on whole sample images, which every form keeps in device memory anyway (plain:
 * generations of at most queue_capacity (pixel, sample) items, 16M by default; wavefront: one
 * image per generation sample). stats may be NULL. */
int x3d_render(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
               uint32_t height, uint32_t spp, int32_t firefly_filter, float *rgba,
               x3d_render_stats *stats);

/* Same, leaving the film in device memory (d_rgba: float[4*w*h] device pointer). */
int x3d_render_device(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera,
                      uint32_t width, uint32_t height, uint32_t spp, float *d_rgba,
                      x3d_render_stats *stats);

/* Tracer-level entry (replaces if_path_tracer::compute_first_hit + sample, pathtracer.h:44-63,
 * pathtracer.cpp:456-472, as tst_pathtracer.cpp:58-86 drives them): radiance estimates for the
 * caller's rays instead of a camera's. rays: float[6*n] (origin, direction; the direction is used
 * as given, as e8util::ray does). The first hits are computed once per call (closest hit, a hit
 * seen from behind dropped, the hit's light recorded), then samples [sample_begin,
 * sample_begin + num_samples) are drawn for every ray with the renderer's tracer and options:
 * sample s of ray k uses PCG stream (seed, k, s), so calls with consecutive sample ranges
 * continue one sequence. radiance: float[3 * n * num_samples], sample-major (sample
 * sample_begin + j of ray k at 3 * (j * n + k)). Always the plain form, no firefly filter.
 * stats may be NULL. */
int x3d_trace_rays(x3d_renderer *r, const x3d_scene *scene, const float *rays, uint32_t n, uint32_t sample_begin,
                   uint32_t num_samples, float *radiance, x3d_render_stats *stats);
/* Same with device pointers (d_rays: float[6*n], d_radiance: float[3*n*num_samples]). */
int x3d_trace_rays_device(x3d_renderer *r, const x3d_scene *scene, const float *d_rays, uint32_t n,
                          uint32_t sample_begin, uint32_t num_samples, float *d_radiance, x3d_render_stats *stats);

/* No counterpart in the reference: x3d_render_device for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same film as x3d_render_device bit for bit (any tracer; always the plain form, no firefly
 * filter), but the call allocates nothing: its working memory (first hits 48 B, running sums
 * 16 B per pixel, traversal-stack spill, one or more sample images of 16 B per pixel) comes from
 * a fixed 128 MiB device array that the library holds on every GPU it runs on until the process
 * ends, with fewer resident workgroups where the stack spill would not fit. Films of more than
 * about 1.6M pixels (fewer for scenes with very deep trees) are refused with
 * X3D_ERR_INVALID_ARGUMENT before anything is launched. Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (from several threads or renderers) race and
 * must be serialised by the caller. The scene must already be committed (its upload allocates),
 * and the renderer already created. */
int x3d_render_device_noalloc(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                              uint32_t height, uint32_t spp, float *d_rgba, x3d_render_stats *stats);

/* PCG draws of stream (seed, pixel, sample) computed on the device (for RNG parity tests). */
int x3d_pcg_draws(uint64_t seed, uint64_t pixel, uint32_t sample, uint32_t n, uint32_t *out);

/* ---- batched rendering (no reference counterpart: pt_image_renderer::render takes one scene,
 * renderer.h:57-59). The renderer must be a UNIDIRECT_MIS renderer; each job's film is
 * bit-identical to x3d_render of that job, in either form. Jobs may share scenes.
 * WAVEFRONT form: after each job's first-hit pass, the paths of all jobs share the generations of
 * one wavefront render (queues, traversal and shading kernels of every bounce, replayed as one
 * hipGraph per generation): each generation holds whole samples of several jobs' pixels, up to
 * queue_capacity paths (fewest-sample jobs first; a job's samples may span generations, in
 * order). Options used: max_depth, seed, queue_capacity, workgroup_size, stack_depth (sort_key,
 * ray_order and batch_split do not apply).
 * PLAIN form: jobs with fewer than batch_split samples share one persistent launch, captured as a
 * hipGraph and replayed while the batch keeps its shape; larger jobs then run one after another
 * in the wavefront form (fewest samples first). Options used: max_depth, seed, batch_split; for
 * the shared launch workgroup_size (64/128/256), stack_depth (8/16), sort_key (NONE: hand tiles
 * out in job order; otherwise most expensive job first); for the large jobs queue_capacity. ---- */
typedef struct x3d_batch_job {
    const x3d_scene *scene;
    x3d_camera camera;
    uint32_t width, height, spp; /* width, height >= 2; spp >= 1 */
    uint32_t reserved;
    float *rgba;       /* film float[4*w*h]: host memory for x3d_render_batch, device memory for
                          x3d_render_batch_device */
    double latency_ms; /* out: device time from the batch's start to the job's completion */
} x3d_batch_job;

typedef struct x3d_batch_stats {
    uint64_t num_samples;     /* sum of w*h*spp */
    uint64_t closest_queries; /* first hits included */
    uint64_t anyhit_queries;
    uint64_t closest_nodes, closest_tests, anyhit_nodes, anyhit_tests;
    double kernel_ms;         /* device time of the whole batch (HIP events) */
    double total_ms;          /* wall time of the call */
    uint32_t tiles;           /* PLAIN: work tiles of 64 pixels in the shared launch */
    uint32_t alone_jobs;      /* PLAIN: jobs at or above batch_split, rendered on their own */
    uint32_t generations;     /* WAVEFRONT: shared generations */
    uint32_t reserved;
} x3d_batch_stats;

int x3d_render_batch(x3d_renderer *r, uint32_t num_jobs, x3d_batch_job *jobs, x3d_batch_stats *stats);
int x3d_render_batch_device(x3d_renderer *r, uint32_t num_jobs, x3d_batch_job *jobs,
                            x3d_batch_stats *stats);

/* ---- differentiation (no reference counterpart: the reference has no differentiable
 * renderer). Gradient of a film loss with respect to material parameters for the
 * UNIDIRECT_MIS tracer (either form: the backward pass always replays paths one per lane;
 * max_depth <= 8, at most 64 materials). ---- */

/* Forward pass of the pair: the film of x3d_render (firefly filter off). */
int x3d_render_forward(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                       uint32_t height, uint32_t spp, float *rgba, x3d_render_stats *stats);

/* Backward pass: given dl_dfilm = dLoss/dfilm (float[4*w*h], rgb used, same layout as the
 * film), writes grad = dLoss/dparams, float[4*num_materials]: for material m, grad[4m..4m+2] is
 * the derivative with respect to its albedo (OREN_NAYAR, COOK_TORR) and grad[4m+3] with respect
 * to its roughness (COOK_TORR); mixtures get zeros (their operands carry the gradient). Each
 * path is re-traced from its PCG stream (path replay; nothing is kept from the forward pass).
 * The sampled directions, sampling densities and MIS weights are held fixed, so the result is
 * the exact derivative of the film for fixed seeds when the GGX sampling density is held
 * (x3d_scene_set_material_params with update_sampling = 0), and an unbiased estimate of the
 * gradient of the expected film in any case. Visibility does not depend on the parameters (no
 * boundary terms). The sums are reproducible run to run. */
int x3d_render_backward(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera, uint32_t width,
                        uint32_t height, uint32_t spp, const float *dl_dfilm, float *grad,
                        x3d_render_stats *stats);
/* Same with device pointers (d_dl_dfilm: float[4*w*h], d_grad: float[4*num_materials]). */
int x3d_render_backward_device(x3d_renderer *r, const x3d_scene *scene, const x3d_camera *camera,
                               uint32_t width, uint32_t height, uint32_t spp, const float *d_dl_dfilm,
                               float *d_grad, x3d_render_stats *stats);

/* Autotuner for a batch: times every workgroup size x LDS stack depth (x sort key, PLAIN form:
 * 12 candidates; WAVEFRONT form: 6) on the given jobs (their films are written), keeps the
 * fastest in the renderer's options. */
int x3d_renderer_autotune_batch(x3d_renderer *r, uint32_t num_jobs, x3d_batch_job *jobs,
                                x3d_autotune_report *report);

/* ---- film compositors (replace clamp_compositor and aces_compositor, compositor.h:35-58,
 * compositor.cpp:29-98): the linear RGBA film of x3d_render -> the 8-bit frame
 * (e8::pixel = 4 x uint8 r, g, b, a per pixel, frame.h:12; pixel (x, y) at x + y*w as
 * if_frame::operator(), frame.cpp:35-41). Per channel byte = (uint8)(pow(CLAMP(v, 0, 1), 1/2.2)
 * * 255), alpha byte = (uint8)(a * 255) (if_compositor::pixel_of). CLAMP clamps all four
 * components to [0, 1] first. ACES multiplies rgb by the exposure and applies
 * x (2.51 x + 0.03) / (x (2.43 x + 0.59) + 0.14) per component; with auto exposure (the
 * default, as aces_compositor's) the exposure is exp of the mean of
 * log(0.299 r + 0.587 g + 0.114 b + 1e-3) over the pixels with any component farther than 1e-5
 * from 0 (a film with none gives NaN, as the reference's 0/0). The transcendental functions
 * are the library's own, correctly rounded except within about 1e-14 of a rounding boundary,
 * and the log sum is a fixed-order double sum: results do not depend on the device or run.
 * NaN channels give byte 0; values the reference would convert out of range (undefined there)
 * saturate to 0 or 255. ---- */
#define X3D_COMPOSITOR_CLAMP 0
#define X3D_COMPOSITOR_ACES 1

typedef struct x3d_compositor x3d_compositor;

typedef struct x3d_compose_stats {
    float exposure;      /* exposure applied (ACES; 1 for CLAMP) */
    uint32_t counted;    /* pixels in the auto-exposure mean (0 without auto exposure) */
    double exposure_ms;  /* device time of the auto-exposure pass (HIP events) */
    double compose_ms;   /* device time of the tonemap / gamma / 8-bit pass */
    double total_ms;     /* wall time of the call */
} x3d_compose_stats;

int x3d_compositor_create(int32_t kind, x3d_compositor **out);
int x3d_compositor_destroy(x3d_compositor *c);
/* aces_compositor::enable_auto_exposure / exposure(e) (compositor.cpp:64-66); ignored by
 * CLAMP. Defaults: auto exposure on, exposure 0. */
int x3d_compositor_enable_auto_exposure(x3d_compositor *c, int32_t on);
int x3d_compositor_set_exposure(x3d_compositor *c, float exposure);
/* if_compositor::commit (compositor.cpp:41-62): rgba float[4*w*h] -> pixels uint8[4*w*h].
 * w*h must be 1 .. 2^31. stats may be NULL. */
int x3d_compositor_commit(x3d_compositor *c, const float *rgba, uint32_t width, uint32_t height,
                          uint8_t *pixels, x3d_compose_stats *stats);
/* Same with device pointers (e.g. the film of x3d_render_device), on the compositor's own
 * stream; returns when the frame is written. */
int x3d_compositor_commit_device(x3d_compositor *c, const float *d_rgba, uint32_t width,
                                 uint32_t height, uint8_t *d_pixels, x3d_compose_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* X3DPT_H */
