"""MI355X-native path-tracing hot path of DaviesX/x3dcore2, behind the C ABI of libx3dpt.so.

Python is a thin ctypes layer over include/x3dpt.h; every computation runs in the HIP library.
Importing the package fails loudly when libx3dpt.so has not been built (there is no CPU
fallback). Build with ``python -m x3dcore2_amd.build`` or ``__graft_entry__.build()``.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import scenes  # noqa: F401  (re-exported)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libx3dpt.so")

# Tracer types (include/x3dpt.h): pathtracer_factory::pt_type order plus the MIS tracer.
PT_NORMAL, PT_POSITION, PT_DIRECT, PT_UNIDIRECT, PT_UNIDIRECT_LT1, PT_BIDIRECT_LT2, PT_BIDIRECT_MIS = range(7)
PT_UNIDIRECT_MIS = 16
FORM_PLAIN, FORM_WAVEFRONT, FORM_AUTO = 0, 1, 2
SORT_DEFAULT, SORT_MATERIAL, SORT_NONE = 0, 1, 2
RAY_ORDER_DEFAULT, RAY_ORDER_SORTED, RAY_ORDER_APPEND = 0, 1, 2
COMPOSITOR_CLAMP, COMPOSITOR_ACES = 0, 1
OBJ_GEOMETRY, OBJ_LIGHT, OBJ_MATERIAL = 1, 3, 4
TRAVERSAL_CULLED, TRAVERSAL_REFERENCE = 0, 1


class X3DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"x3dpt error {code}: {msg}")
        self.code = code


class MaterialDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("albedo", C.c_float * 3), ("sigma", C.c_float), ("roughness", C.c_float),
                ("ior_re", C.c_float), ("ior_im", C.c_float), ("sub0", C.c_int32), ("sub1", C.c_int32),
                ("ratio", C.c_float)]


class SceneInfo(C.Structure):
    _fields_ = [("num_meshes", C.c_uint32), ("num_tris", C.c_uint32), ("num_nodes", C.c_uint32),
                ("num_interior", C.c_uint32), ("num_lights", C.c_uint32), ("num_materials", C.c_uint32),
                ("max_leaf_depth", C.c_uint32), ("device_bytes", C.c_uint64), ("build_seconds", C.c_double),
                ("upload_seconds", C.c_double)]


class Hit(C.Structure):
    _fields_ = [("geo", C.c_int32), ("tri", C.c_int32), ("prim", C.c_int32), ("t", C.c_float), ("b1", C.c_float),
                ("b2", C.c_float), ("vertex", C.c_float * 3), ("normal", C.c_float * 3)]


HIT_DTYPE = np.dtype([("geo", np.int32), ("tri", np.int32), ("prim", np.int32), ("t", np.float32),
                      ("b1", np.float32), ("b2", np.float32), ("vertex", np.float32, 3),
                      ("normal", np.float32, 3)])
assert HIT_DTYPE.itemsize == C.sizeof(Hit)


class ObjMesh(C.Structure):
    _fields_ = [("num_verts", C.c_uint32), ("num_tris", C.c_uint32), ("verts", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)), ("tris", C.POINTER(C.c_uint32))]


class CameraDesc(C.Structure):
    _fields_ = [("sensor_size", C.c_float), ("focal_length", C.c_float), ("aspect", C.c_float),
                ("transform", C.c_float * 16)]


class RenderOptions(C.Structure):
    _fields_ = [("pt_type", C.c_int32), ("form", C.c_int32), ("max_depth", C.c_uint32),
                ("max_path_len", C.c_uint32), ("seed", C.c_uint64), ("workgroup_size", C.c_uint32),
                ("queue_capacity", C.c_uint32), ("sort_key", C.c_int32), ("stack_depth", C.c_uint32),
                ("profile", C.c_int32), ("batch_split", C.c_uint32), ("ray_order", C.c_int32),
                ("reserved", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("num_samples", C.c_uint64), ("closest_queries", C.c_uint64), ("anyhit_queries", C.c_uint64),
                ("closest_nodes", C.c_uint64), ("closest_tests", C.c_uint64), ("anyhit_nodes", C.c_uint64),
                ("anyhit_tests", C.c_uint64), ("first_hit_nodes", C.c_uint64), ("first_hit_tests", C.c_uint64),
                ("kernel_ms", C.c_double), ("first_hit_ms", C.c_double),
                ("trace_ms", C.c_double), ("total_ms", C.c_double), ("extend_ms", C.c_double),
                ("shadow_ms", C.c_double), ("trace_launches", C.c_uint32), ("form", C.c_int32)]


class AutotuneReport(C.Structure):
    _fields_ = [("workgroup_size", C.c_uint32), ("stack_depth", C.c_uint32), ("queue_capacity", C.c_uint32),
                ("sort_key", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("candidates", C.c_uint32), ("reserved", C.c_uint32)]


_lib: Optional[C.CDLL] = None

class BatchJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", CameraDesc), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("reserved", C.c_uint32), ("rgba", C.c_void_p), ("latency_ms", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("num_samples", C.c_uint64), ("closest_queries", C.c_uint64), ("anyhit_queries", C.c_uint64),
                ("closest_nodes", C.c_uint64), ("closest_tests", C.c_uint64), ("anyhit_nodes", C.c_uint64),
                ("anyhit_tests", C.c_uint64), ("kernel_ms", C.c_double), ("total_ms", C.c_double),
                ("tiles", C.c_uint32), ("alone_jobs", C.c_uint32), ("generations", C.c_uint32),
                ("reserved", C.c_uint32)]


class ComposeStats(C.Structure):
    _fields_ = [("exposure", C.c_float), ("counted", C.c_uint32), ("exposure_ms", C.c_double),
                ("compose_ms", C.c_double), ("total_ms", C.c_double)]


# Every exported symbol declared in include/x3dpt.h.
EXPORTS = [
    "x3d_last_error", "x3d_version", "x3d_scene_create", "x3d_scene_destroy", "x3d_scene_add_material",
    "x3d_scene_add_mesh", "x3d_scene_add_area_light", "x3d_scene_add_sky_light", "x3d_scene_commit",
    "x3d_scene_get_info",
    "x3d_scene_dump_bvh", "x3d_intersect", "x3d_occluded", "x3d_intersect_device", "x3d_occluded_device",
    "x3d_bsdf_eval", "x3d_bsdf_pdf", "x3d_bsdf_sample", "x3d_camera_rays", "x3d_render_options_default",
    "x3d_renderer_create", "x3d_renderer_destroy", "x3d_render", "x3d_render_device", "x3d_pcg_draws",
    "x3d_trace_rays", "x3d_trace_rays_device", "x3d_render_device_noalloc",
    "x3d_renderer_autotune", "x3d_render_batch", "x3d_render_batch_device", "x3d_renderer_autotune_batch",
    "x3d_scene_set_material_params", "x3d_render_forward", "x3d_render_backward", "x3d_render_backward_device",
    "x3d_obj_load", "x3d_obj_free", "x3d_scene_add_obj",
    "x3d_compositor_create", "x3d_compositor_destroy", "x3d_compositor_enable_auto_exposure",
    "x3d_compositor_set_exposure", "x3d_compositor_commit", "x3d_compositor_commit_device",
    "x3d_scene_load_geometry", "x3d_scene_load_material", "x3d_scene_load_light", "x3d_scene_load_sky_light",
    "x3d_scene_unload",
    "x3d_scene_geometry_obj_id", "x3d_transform_mesh",
]


def lib() -> C.CDLL:
    """Loads libx3dpt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m x3dcore2_amd.build` "
                          "(the renderer has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, u32, i32, u64 = C.c_void_p, C.c_uint32, C.c_int32, C.c_uint64
    fp = C.POINTER(C.c_float)
    L.x3d_last_error.restype = C.c_char_p
    L.x3d_version.restype = C.c_char_p
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("x3d_last_error", "x3d_version"):
            f.restype = C.c_int
    L.x3d_scene_create.argtypes = [C.POINTER(vp)]
    L.x3d_scene_destroy.argtypes = [vp]
    L.x3d_scene_add_material.argtypes = [vp, C.POINTER(MaterialDesc), C.POINTER(i32)]
    L.x3d_scene_add_mesh.argtypes = [vp, vp, vp, vp, u32, vp, u32, i32, C.POINTER(i32)]
    L.x3d_obj_load.argtypes = [C.c_char_p, C.POINTER(ObjMesh)]
    L.x3d_obj_free.argtypes = [C.POINTER(ObjMesh)]
    L.x3d_scene_add_obj.argtypes = [vp, C.c_char_p, i32, C.POINTER(i32)]
    L.x3d_scene_add_area_light.argtypes = [vp, i32, fp, C.POINTER(i32)]
    L.x3d_scene_add_sky_light.argtypes = [vp, fp, C.POINTER(i32)]
    L.x3d_scene_commit.argtypes = [vp]
    L.x3d_scene_get_info.argtypes = [vp, C.POINTER(SceneInfo)]
    L.x3d_scene_dump_bvh.argtypes = [vp, vp, vp, vp]
    L.x3d_intersect.argtypes = [vp, vp, u32, i32, vp]
    L.x3d_occluded.argtypes = [vp, vp, vp, vp, u32, vp]
    L.x3d_intersect_device.argtypes = [vp, vp, u32, i32, vp, vp]
    L.x3d_occluded_device.argtypes = [vp, vp, vp, vp, u32, vp, vp]
    L.x3d_bsdf_eval.argtypes = [vp, i32, u32, vp, vp, vp, vp]
    L.x3d_bsdf_pdf.argtypes = [vp, i32, u32, vp, vp, vp, vp]
    L.x3d_bsdf_sample.argtypes = [vp, i32, u64, u32, vp, vp, i32, vp, vp]
    L.x3d_camera_rays.argtypes = [C.POINTER(CameraDesc), u32, u32, vp]
    L.x3d_render_options_default.argtypes = [i32, C.POINTER(RenderOptions)]
    L.x3d_renderer_create.argtypes = [C.POINTER(RenderOptions), C.POINTER(vp)]
    L.x3d_renderer_destroy.argtypes = [vp]
    L.x3d_render.argtypes = [vp, vp, C.POINTER(CameraDesc), u32, u32, u32, i32, vp, C.POINTER(RenderStats)]
    L.x3d_render_device.argtypes = [vp, vp, C.POINTER(CameraDesc), u32, u32, u32, vp, C.POINTER(RenderStats)]
    L.x3d_pcg_draws.argtypes = [u64, u64, u32, u32, vp]
    L.x3d_render_device_noalloc.argtypes = [vp, vp, C.POINTER(CameraDesc), u32, u32, u32, vp,
                                            C.POINTER(RenderStats)]
    L.x3d_trace_rays.argtypes = [vp, vp, vp, u32, u32, u32, vp, C.POINTER(RenderStats)]
    L.x3d_trace_rays_device.argtypes = [vp, vp, vp, u32, u32, u32, vp, C.POINTER(RenderStats)]
    L.x3d_renderer_autotune.argtypes = [vp, vp, C.POINTER(CameraDesc), u32, u32, u32, C.POINTER(AutotuneReport)]
    L.x3d_render_batch.argtypes = [vp, u32, C.POINTER(BatchJob), C.POINTER(BatchStats)]
    L.x3d_render_batch_device.argtypes = [vp, u32, C.POINTER(BatchJob), C.POINTER(BatchStats)]
    L.x3d_renderer_autotune_batch.argtypes = [vp, u32, C.POINTER(BatchJob), C.POINTER(AutotuneReport)]
    L.x3d_scene_set_material_params.argtypes = [vp, u32, u32, vp, vp, i32]
    L.x3d_render_forward.argtypes = [vp, vp, C.POINTER(CameraDesc), u32, u32, u32, vp, C.POINTER(RenderStats)]
    L.x3d_render_backward.argtypes = [vp, vp, C.POINTER(CameraDesc), u32, u32, u32, vp, vp, C.POINTER(RenderStats)]
    L.x3d_render_backward_device.argtypes = [vp, vp, C.POINTER(CameraDesc), u32, u32, u32, vp, vp,
                                             C.POINTER(RenderStats)]
    L.x3d_compositor_create.argtypes = [i32, C.POINTER(vp)]
    L.x3d_compositor_destroy.argtypes = [vp]
    L.x3d_compositor_enable_auto_exposure.argtypes = [vp, i32]
    L.x3d_compositor_set_exposure.argtypes = [vp, C.c_float]
    L.x3d_compositor_commit.argtypes = [vp, vp, u32, u32, vp, C.POINTER(ComposeStats)]
    L.x3d_compositor_commit_device.argtypes = [vp, vp, u32, u32, vp, C.POINTER(ComposeStats)]
    L.x3d_scene_load_geometry.argtypes = [vp, u64, vp, vp, vp, u32, vp, u32, i32, u64, vp]
    L.x3d_scene_load_material.argtypes = [vp, u64, C.POINTER(MaterialDesc)]
    L.x3d_scene_load_light.argtypes = [vp, u64, u64, fp]
    L.x3d_scene_load_sky_light.argtypes = [vp, u64, fp]
    L.x3d_scene_unload.argtypes = [vp, i32, u64]
    L.x3d_scene_geometry_obj_id.argtypes = [vp, i32, C.POINTER(u64)]
    L.x3d_transform_mesh.argtypes = [vp, u32, vp, vp, vp, vp]
    missing = [n for n in EXPORTS if n not in ("x3d_last_error", "x3d_version") and getattr(L, n).argtypes is None]
    if missing:  # a pointer passed without argtypes would be truncated to a C int
        raise ImportError(f"no ctypes signature for {missing}")
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise X3DError(rc, lib().x3d_last_error().decode())


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def camera_desc(cam) -> CameraDesc:
    c = CameraDesc()
    c.sensor_size, c.focal_length, c.aspect = cam.sensor, cam.focal, cam.aspect
    for k in range(16):
        c.transform[k] = float(cam.transform[k])
    return c


class Scene:
    """A committed scene resident in device memory (x3d_scene)."""

    def __init__(self, desc=None, commit: bool = True):
        L = lib()
        h = C.c_void_p()
        _check(L.x3d_scene_create(C.byref(h)))
        self.h = h
        self.desc = desc
        self._keep = []
        if desc is not None:
            self.load(desc)
            if commit:
                self.commit()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.x3d_scene_destroy(self.h)
            self.h = None

    @staticmethod
    def _desc(m) -> MaterialDesc:
        d = MaterialDesc()
        d.kind = m.kind
        for k in range(3):
            d.albedo[k] = m.albedo[k]
        d.sigma, d.roughness = m.sigma, m.roughness
        d.ior_re, d.ior_im = m.ior
        d.sub0, d.sub1, d.ratio = m.sub0, m.sub1, m.ratio
        return d

    # ---- objects by id (the if_obj_actuator protocol: load / unload / commit) ----
    def load_geometry(self, obj_id, verts, normals, tris, material=None, transform=None):
        """material: obj id of a material (None: the fail-safe); transform: 4x4 column-major
        (16 floats) or None."""
        v, n = _f32(verts), _f32(normals)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        m = None if transform is None else _f32(transform).reshape(-1)
        _check(lib().x3d_scene_load_geometry(self.h, int(obj_id), _ptr(v), _ptr(n), None, v.shape[0], _ptr(t),
                                             t.shape[0], 0 if material is None else 1,
                                             0 if material is None else int(material), _ptr(m)))

    def load_material(self, obj_id, m):
        _check(lib().x3d_scene_load_material(self.h, int(obj_id), C.byref(self._desc(m))))

    def load_light(self, obj_id, geometry_obj_id, radiance):
        r = (C.c_float * 3)(*radiance)
        _check(lib().x3d_scene_load_light(self.h, int(obj_id), int(geometry_obj_id), r))

    def load_sky_light(self, obj_id, radiance):
        r = (C.c_float * 3)(*radiance)
        _check(lib().x3d_scene_load_sky_light(self.h, int(obj_id), r))

    def unload(self, protocol, obj_id):
        _check(lib().x3d_scene_unload(self.h, int(protocol), int(obj_id)))

    def geometry_obj_id(self, geo) -> int:
        out = C.c_uint64()
        _check(lib().x3d_scene_geometry_obj_id(self.h, int(geo), C.byref(out)))
        return out.value

    def add_material(self, m) -> int:
        d = MaterialDesc()
        d.kind = m.kind
        for k in range(3):
            d.albedo[k] = m.albedo[k]
        d.sigma, d.roughness = m.sigma, m.roughness
        d.ior_re, d.ior_im = m.ior
        d.sub0, d.sub1, d.ratio = m.sub0, m.sub1, m.ratio
        out = C.c_int32()
        _check(lib().x3d_scene_add_material(self.h, C.byref(d), C.byref(out)))
        return out.value

    def add_mesh(self, verts, normals, tris, material=-1) -> int:
        v, n = _f32(verts), _f32(normals)
        t = np.ascontiguousarray(tris, dtype=np.uint32)
        out = C.c_int32()
        _check(lib().x3d_scene_add_mesh(self.h, _ptr(v), _ptr(n), None, v.shape[0], _ptr(t), t.shape[0],
                                        int(material), C.byref(out)))
        return out.value

    def add_obj(self, path, material=-1) -> int:
        """x3d_scene_add_obj: a Wavefront OBJ mesh with the reference loader's rules."""
        out = C.c_int32()
        _check(lib().x3d_scene_add_obj(self.h, os.fsencode(path), int(material), C.byref(out)))
        return out.value

    def add_area_light(self, mesh_id, radiance) -> int:
        r = (C.c_float * 3)(*radiance)
        out = C.c_int32()
        _check(lib().x3d_scene_add_area_light(self.h, int(mesh_id), r, C.byref(out)))
        return out.value

    def add_sky_light(self, radiance) -> int:
        r = (C.c_float * 3)(*radiance)
        out = C.c_int32()
        _check(lib().x3d_scene_add_sky_light(self.h, r, C.byref(out)))
        return out.value

    def load(self, desc):
        """Loads a scenes.SceneDesc; a light entry with mesh index < 0 is the sky light."""
        ids = [self.add_material(m) for m in desc.materials]
        for mesh in desc.meshes:
            self.add_mesh(mesh.verts, mesh.normals, mesh.tris, ids[mesh.material] if mesh.material >= 0 else -1)
        for mi, rad in desc.lights:
            if mi < 0:
                self.add_sky_light(rad)
            else:
                self.add_area_light(mi, rad)

    def commit(self):
        _check(lib().x3d_scene_commit(self.h))

    def info(self) -> SceneInfo:
        i = SceneInfo()
        _check(lib().x3d_scene_get_info(self.h, C.byref(i)))
        return i

    @property
    def num_materials(self) -> int:
        """Materials added by the caller (the committed table also holds the fail-safe one)."""
        return self.info().num_materials - 1

    def set_material_params(self, first, albedo=None, roughness=None, update_sampling=True):
        """albedo: [count, 3] or None; roughness: [count] or None (x3d_scene_set_material_params)."""
        a = None if albedo is None else np.ascontiguousarray(albedo, np.float32).reshape(-1, 3)
        r = None if roughness is None else np.ascontiguousarray(roughness, np.float32).reshape(-1)
        count = a.shape[0] if a is not None else r.shape[0]
        _check(lib().x3d_scene_set_material_params(self.h, first, count, _ptr(a) if a is not None else None,
                                                   _ptr(r) if r is not None else None, int(update_sampling)))

    def bvh(self):
        i = self.info()
        bounds = np.zeros((i.num_nodes, 6), np.float32)
        meta = np.zeros((i.num_nodes, 4), np.uint32)
        prims = np.zeros((i.num_tris, 2), np.int32)
        _check(lib().x3d_scene_dump_bvh(self.h, _ptr(bounds), _ptr(meta), _ptr(prims)))
        return bounds, meta, prims

    def intersect(self, rays, traversal=TRAVERSAL_CULLED) -> np.ndarray:
        r = _f32(rays)
        hits = np.zeros(r.shape[0], HIT_DTYPE)
        _check(lib().x3d_intersect(self.h, _ptr(r), r.shape[0], traversal, _ptr(hits)))
        return hits

    def occluded(self, rays, tmin, tmax) -> np.ndarray:
        r = _f32(rays)
        out = np.zeros(r.shape[0], np.uint8)
        _check(lib().x3d_occluded(self.h, _ptr(r), _ptr(_f32(tmin)), _ptr(_f32(tmax)), r.shape[0], _ptr(out)))
        return out

    def bsdf_eval(self, mat, nrm, o, i) -> np.ndarray:
        n = nrm.shape[0]
        f = np.zeros((n, 3), np.float32)
        _check(lib().x3d_bsdf_eval(self.h, mat, n, _ptr(_f32(nrm)), _ptr(_f32(o)), _ptr(_f32(i)), _ptr(f)))
        return f

    def bsdf_pdf(self, mat, nrm, o, i) -> np.ndarray:
        n = nrm.shape[0]
        p = np.zeros(n, np.float32)
        _check(lib().x3d_bsdf_pdf(self.h, mat, n, _ptr(_f32(nrm)), _ptr(_f32(o)), _ptr(_f32(i)), _ptr(p)))
        return p

    def bsdf_sample(self, mat, seed, nrm, o, reference_mixture=False):
        n = nrm.shape[0]
        i = np.zeros((n, 3), np.float32)
        d = np.zeros(n, np.float32)
        _check(lib().x3d_bsdf_sample(self.h, mat, seed, n, _ptr(_f32(nrm)), _ptr(_f32(o)), int(reference_mixture),
                                     _ptr(i), _ptr(d)))
        return i, d


class Renderer:
    """pt_image_renderer equivalent (x3d_renderer). `form` defaults to FORM_PLAIN here so that a
    test names the form it exercises; x3d_render_options_default (the C default) gives FORM_AUTO
    to the tracers that have a wavefront form. Every form gives the same film."""

    def __init__(self, pt_type=PT_UNIDIRECT_MIS, form=FORM_PLAIN, max_depth=5, max_path_len=8, seed=1,
                 workgroup_size=0, queue_capacity=0, sort_key=0, stack_depth=0, profile=0, batch_split=0,
                 ray_order=0):
        L = lib()
        o = RenderOptions()
        _check(L.x3d_render_options_default(pt_type, C.byref(o)))
        o.form, o.max_depth, o.max_path_len, o.seed = form, max_depth, max_path_len, seed
        o.workgroup_size, o.queue_capacity, o.sort_key = workgroup_size, queue_capacity, sort_key
        o.stack_depth, o.profile, o.batch_split, o.ray_order = stack_depth, profile, batch_split, ray_order
        h = C.c_void_p()
        _check(L.x3d_renderer_create(C.byref(o), C.byref(h)))
        self.h = h
        self.opts = o

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.x3d_renderer_destroy(self.h)
            self.h = None

    def render(self, scene: Scene, camera, width, height, spp, firefly_filter=False):
        """Film [h*w, 4] (rgb = mean radiance, alpha = 1/spp) and RenderStats. firefly_filter
        caps isolated bright samples in each sample image before the per-pixel sum."""
        rgba = np.zeros((height * width, 4), np.float32)
        st = RenderStats()
        _check(lib().x3d_render(self.h, scene.h, C.byref(camera_desc(camera)), width, height, spp,
                                1 if firefly_filter else 0, _ptr(rgba),
                                C.byref(st)))
        return rgba, st

    def trace_rays(self, scene: Scene, rays, num_samples, sample_begin=0):
        """x3d_trace_rays: radiance estimates for caller rays (float32 [n, 6]: origin, direction).
        Returns ([num_samples, n, 3] float32, RenderStats)."""
        r = _f32(rays).reshape(-1, 6)
        out = np.zeros((num_samples, r.shape[0], 3), np.float32)
        st = RenderStats()
        _check(lib().x3d_trace_rays(self.h, scene.h, _ptr(r), r.shape[0], sample_begin, num_samples, _ptr(out),
                                    C.byref(st)))
        return out, st

    def render_device(self, scene: Scene, camera, width, height, spp, d_rgba: int):
        st = RenderStats()
        _check(lib().x3d_render_device(self.h, scene.h, C.byref(camera_desc(camera)), width, height, spp,
                                       C.c_void_p(d_rgba), C.byref(st)))
        return st

    @staticmethod
    def _jobs(items, films):
        arr = (BatchJob * len(items))()
        for k, (scene, camera, w, h, spp) in enumerate(items):
            j = arr[k]
            j.scene, j.camera, j.width, j.height, j.spp = scene.h.value, camera_desc(camera), w, h, spp
            j.rgba = films[k] if isinstance(films[k], int) else films[k].ctypes.data
        return arr

    def render_batch(self, items):
        """items: [(Scene, camera, width, height, spp)]. Returns (films [h*w, 4] float32 per job,
        BatchStats, per-job latency in ms)."""
        films = [np.zeros((h * w, 4), np.float32) for (_, _, w, h, _) in items]
        arr = self._jobs(items, films)
        st = BatchStats()
        _check(lib().x3d_render_batch(self.h, len(items), arr, C.byref(st)))
        return films, st, [arr[k].latency_ms for k in range(len(items))]

    def render_batch_device(self, items, d_films):
        """Same with device films (device pointers as ints)."""
        arr = self._jobs(items, d_films)
        st = BatchStats()
        _check(lib().x3d_render_batch_device(self.h, len(items), arr, C.byref(st)))
        return st, [arr[k].latency_ms for k in range(len(items))]

    def autotune_batch(self, items) -> AutotuneReport:
        films = [np.zeros((h * w, 4), np.float32) for (_, _, w, h, _) in items]
        arr = self._jobs(items, films)
        rep = AutotuneReport()
        _check(lib().x3d_renderer_autotune_batch(self.h, len(items), arr, C.byref(rep)))
        return rep

    def forward(self, scene: Scene, camera, width, height, spp):
        rgba = np.zeros((height * width, 4), np.float32)
        st = RenderStats()
        _check(lib().x3d_render_forward(self.h, scene.h, C.byref(camera_desc(camera)), width, height, spp, _ptr(rgba),
                                        C.byref(st)))
        return rgba, st

    def backward(self, scene: Scene, camera, width, height, spp, dl_dfilm):
        """dl_dfilm: [h*w, 4] (rgb used). Returns ([num_materials, 4] float32 gradient:
        albedo rgb, roughness), stats."""
        dl = np.ascontiguousarray(dl_dfilm, np.float32).reshape(height * width, 4)
        grad = np.zeros((scene.num_materials, 4), np.float32)
        st = RenderStats()
        _check(lib().x3d_render_backward(self.h, scene.h, C.byref(camera_desc(camera)), width, height, spp,
                                         _ptr(dl), _ptr(grad), C.byref(st)))
        return grad, st

    def autotune(self, scene: Scene, camera, width, height, probe_spp) -> AutotuneReport:
        rep = AutotuneReport()
        _check(lib().x3d_renderer_autotune(self.h, scene.h, C.byref(camera_desc(camera)), width, height, probe_spp,
                                           C.byref(rep)))
        return rep


def load_obj(path):
    """x3d_obj_load: (verts [n,3], normals [n,3], uvs [n,2], tris [m,3]) of a Wavefront OBJ mesh."""
    m = ObjMesh()
    _check(lib().x3d_obj_load(os.fsencode(path), C.byref(m)))
    try:
        nv, nt = m.num_verts, m.num_tris
        out = (np.ctypeslib.as_array(m.verts, (nv, 3)).copy(), np.ctypeslib.as_array(m.normals, (nv, 3)).copy(),
               np.ctypeslib.as_array(m.uvs, (nv, 2)).copy(),
               np.ctypeslib.as_array(m.tris, (max(nt, 1), 3))[:nt].copy())
    finally:
        lib().x3d_obj_free(C.byref(m))
    return out


def camera_rays(camera, width, height) -> np.ndarray:
    rays = np.zeros((width * height, 6), np.float32)
    _check(lib().x3d_camera_rays(C.byref(camera_desc(camera)), width, height, _ptr(rays)))
    return rays


def pcg_draws(seed, pixel, sample, n) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    _check(lib().x3d_pcg_draws(seed, pixel, sample, n, _ptr(out)))
    return out


class Compositor:
    """clamp_compositor / aces_compositor equivalent (x3d_compositor): linear film -> 8-bit
    RGBA frame on the GPU. ACES defaults to auto exposure, as the reference's."""

    def __init__(self, kind=COMPOSITOR_ACES):
        h = C.c_void_p()
        _check(lib().x3d_compositor_create(int(kind), C.byref(h)))
        self.h = h
        self.kind = kind

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.x3d_compositor_destroy(self.h)
            self.h = None

    def enable_auto_exposure(self, on=True):
        _check(lib().x3d_compositor_enable_auto_exposure(self.h, 1 if on else 0))

    def set_exposure(self, e: float):
        _check(lib().x3d_compositor_set_exposure(self.h, float(e)))

    def commit(self, film: np.ndarray, width: int, height: int):
        """film [h*w, 4] float32 -> (frame uint8 [h*w, 4] r, g, b, a, ComposeStats)."""
        f = _f32(film)
        if f.size != 4 * width * height:
            raise ValueError("film size does not match width x height")
        frame = np.zeros((width * height, 4), np.uint8)
        st = ComposeStats()
        _check(lib().x3d_compositor_commit(self.h, _ptr(f), width, height, _ptr(frame), C.byref(st)))
        return frame, st

    def commit_device(self, d_film: int, width: int, height: int, d_frame: int):
        st = ComposeStats()
        _check(lib().x3d_compositor_commit_device(self.h, C.c_void_p(d_film), width, height, C.c_void_p(d_frame),
                                                  C.byref(st)))
        return st


def transform_mesh(transform, verts, normals):
    """x3d_transform_mesh: trimesh::transform of (verts, normals) by a column-major 4x4."""
    m = _f32(transform).reshape(-1)
    v, n = _f32(verts).reshape(-1, 3), _f32(normals).reshape(-1, 3)
    ov, on = np.zeros_like(v), np.zeros_like(n)
    _check(lib().x3d_transform_mesh(_ptr(m), v.shape[0], _ptr(v), _ptr(n), _ptr(ov), _ptr(on)))
    return ov, on
