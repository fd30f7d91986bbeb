"""ctypes bindings for the CPU checkers (test infrastructure only).

`Oracle` wraps oracle/liboracle.so (the C++ restatement); `Reference` wraps
oracle/_ref/libx3dref.so (the reference's own sources built by `make -C oracle ref`). Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use this module.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libx3dref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")

TRAV_REF, TRAV_CULLED, TRAV_BRUTE = 0, 1, 2
# Tracer ids: pathtracer_factory::pt_type order (corelib/src/pathtracerfact.h:12) plus the
# config-mode MIS tracer.
TR_NORMAL, TR_POSITION, TR_DIRECT, TR_UNIDIRECT, TR_UNIDIRECT_LT1, TR_BIDIRECT_LT2, TR_BIDIRECT_MIS = range(7)
TR_UNIDIRECT_MIS = 16


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def oracle_available() -> bool:
    return os.path.exists(ORACLE_SO)


def reference_available() -> bool:
    return os.path.exists(REF_SO)


class _Backend:
    prefix = ""

    def _load_desc(self, desc):
        p = self.prefix
        lib = self.lib
        mat_ids = []
        for m in desc.materials:
            if m.kind == 0:
                mat_ids.append(getattr(lib, p + "add_oren_nayar")(self.h, _f32(m.albedo), C.c_float(m.sigma)))
            elif m.kind == 1:
                mat_ids.append(getattr(lib, p + "add_cook_torr")(self.h, _f32(m.albedo), C.c_float(m.roughness),
                                                                 C.c_float(m.ior[0]), C.c_float(m.ior[1])))
            else:
                mat_ids.append(getattr(lib, p + "add_mixture")(self.h, mat_ids[m.sub0], mat_ids[m.sub1],
                                                               C.c_float(m.ratio)))
        for mesh in desc.meshes:
            getattr(lib, p + "add_mesh")(self.h, _f32(mesh.verts), _f32(mesh.normals), mesh.verts.shape[0],
                                         np.ascontiguousarray(mesh.tris, np.uint32), mesh.tris.shape[0],
                                         mat_ids[mesh.material] if mesh.material >= 0 else -1)
        for mi, rad in desc.lights:
            if mi < 0:  # the sky
                getattr(lib, p + "add_sky_light")(self.h, _f32(rad))
            else:
                getattr(lib, p + "add_area_light")(self.h, mi, _f32(rad))
        getattr(lib, p + "commit")(self.h)

    def light_info(self, light):
        """(power [3], sky reference point [3], sky diameter) after commit."""
        out = np.zeros(7, np.float32)
        getattr(self.lib, self.prefix + "light_info")(self.h, light, out)
        return out[:3], out[3:6], out[6]

    def light_eval(self, light, l, nl, nt):
        """if_light::eval of light `light` at connections l (target - light point), light normals
        nl and target normals nt ([n, 3] each) -> [n, 3] float32."""
        l, nl, nt = _f32(l), _f32(nl), _f32(nt)
        out = np.zeros_like(l)
        getattr(self.lib, self.prefix + "light_eval")(self.h, light, l.shape[0], l, nl, nt, out)
        return out

    def bvh(self):
        nn, np_ = C.c_uint32(), C.c_uint32()
        getattr(self.lib, self.prefix + "bvh_size")(self.h, C.byref(nn), C.byref(np_))
        bounds = np.zeros((nn.value, 6), np.float32)
        meta = np.zeros((nn.value, 4), np.uint32)
        prims = np.zeros((np_.value, 2), np.int32)
        getattr(self.lib, self.prefix + "bvh_dump")(self.h, bounds, meta, prims)
        return bounds, meta, prims


class Oracle(_Backend):
    prefix = "orc_"

    def __init__(self, desc=None):
        lib = C.CDLL(ORACLE_SO)
        self.lib = lib
        lib.orc_scene_new.restype = C.c_void_p
        lib.orc_scene_free.argtypes = [C.c_void_p]
        lib.orc_add_oren_nayar.argtypes = [C.c_void_p, _f32p, C.c_float]
        lib.orc_add_cook_torr.argtypes = [C.c_void_p, _f32p, C.c_float, C.c_float, C.c_float]
        lib.orc_add_mixture.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float]
        lib.orc_add_mesh.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, _u32p, C.c_uint32, C.c_int]
        lib.orc_add_area_light.argtypes = [C.c_void_p, C.c_int, _f32p]
        lib.orc_add_sky_light.argtypes = [C.c_void_p, _f32p]
        lib.orc_light_info.argtypes = [C.c_void_p, C.c_int, _f32p]
        lib.orc_light_eval.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _f32p, _f32p, _f32p, _f32p]
        lib.orc_commit.argtypes = [C.c_void_p]
        lib.orc_bvh_size.argtypes = [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.orc_bvh_dump.argtypes = [C.c_void_p, _f32p, _u32p, _i32p]
        lib.orc_intersect.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_int, _i32p, _i32p, _i32p, _f32p, _f32p,
                                      _f32p, _f32p, C.c_void_p]
        lib.orc_occluded.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, C.c_uint32, C.c_int, _u8p, C.c_void_p]
        lib.orc_bsdf_eval.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _f32p, _f32p, _f32p, C.c_int, _f64p]
        lib.orc_bsdf_pdf.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _f32p, _f32p, _f32p, C.c_int, _f64p]
        lib.orc_bsdf_sample.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint32, _f32p, _f32p, C.c_int,
                                        C.c_int, _f64p, _f64p]
        lib.orc_bsdf_sample_u.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _f32p, _f32p, _f32p, C.c_uint32,
                                          C.c_int, _f64p, _f64p]
        lib.orc_pcg_draws.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _u32p]
        lib.orc_sincos_2pi.argtypes = [C.c_uint32, _f32p, _f32p, _f32p]
        lib.orc_camera_rays.argtypes = [C.c_float, C.c_float, C.c_float, _f32p, C.c_uint32, C.c_uint32, _f32p]
        lib.orc_render.argtypes = [C.c_void_p, _f32p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint64, C.c_int,
                                   C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_int,
                                   C.c_uint32, C.c_void_p, C.c_void_p]
        lib.orc_set_material_params.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                                C.c_int]
        lib.orc_render_backward.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.c_uint32, _f32p, C.c_int, C.c_uint32, _f64p]
        lib.orc_firefly_accumulate.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        lib.orc_trace_rays.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_uint64,
                                       C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        lib.orc_render_firefly.argtypes = [C.c_void_p, _f32p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint64,
                                           C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        self.h = C.c_void_p(lib.orc_scene_new())
        self.desc = desc
        if desc is not None:
            self._load_desc(desc)

    def set_material_params(self, first, albedo=None, roughness=None, update_sampling=True):
        """albedo: [count, 3] or None; roughness: [count] or None."""
        a = None if albedo is None else np.ascontiguousarray(albedo, np.float32).reshape(-1, 3)
        r = None if roughness is None else np.ascontiguousarray(roughness, np.float32).reshape(-1)
        count = a.shape[0] if a is not None else r.shape[0]
        self.lib.orc_set_material_params(self.h, first, count, a.ctypes.data if a is not None else None,
                                         r.ctypes.data if r is not None else None, int(update_sampling))

    def firefly_accumulate(self, samples):
        """samples: [spp, h, w, 3] float32 -> per-pixel in-order sums [h, w, 3] of the
        firefly-filtered samples."""
        x = np.ascontiguousarray(samples, np.float32)
        spp, h, w, _ = x.shape
        out = np.zeros((h, w, 3), np.float32)
        self.lib.orc_firefly_accumulate(x, w, h, spp, out)
        return out

    def trace_rays(self, tracer, rays, num_samples, sample_begin=0, seed=1, max_depth=5, max_path_len=8,
                   threads=None):
        """Radiance of samples [sample_begin, sample_begin + num_samples) of caller rays
        (float32 [n, 6]), sample s of ray k on PCG stream (seed, k, s): [num_samples, n, 3]."""
        r = _f32(rays).reshape(-1, 6)
        out = np.zeros((num_samples, r.shape[0], 3), np.float32)
        self.lib.orc_trace_rays(self.h, r.reshape(-1), r.shape[0], tracer, max_depth, max_path_len, seed,
                                sample_begin, num_samples, threads or os.cpu_count() or 1, out.reshape(-1))
        return out

    def render_firefly(self, tracer, width, height, spp, seed=1, max_depth=5, max_path_len=8, threads=None):
        """orc_render (FP32) with the firefly filter: rgba [h*w, 4] float32."""
        cam = self.desc.camera.packed()
        out = np.zeros((width * height, 4), np.float32)
        self.lib.orc_render_firefly(self.h, cam, tracer, max_depth, max_path_len, seed, width, height, spp,
                                    threads or os.cpu_count() or 1, out)
        return out

    def render_backward(self, width, height, spp, dl_dfilm, seed=1, max_depth=5, prec=1, threads=None,
                        camera=None):
        """Gradient of a film loss (dl_dfilm: [h*w, 4]) w.r.t. (albedo rgb, roughness) of every
        material: [num_materials, 4] float64."""
        cam = (camera or self.desc.camera).packed()
        nm = len(self.desc.materials)
        grad = np.zeros((nm, 4), np.float64)
        self.lib.orc_render_backward(self.h, cam, max_depth, seed, width, height, spp, _f32(dl_dfilm), prec,
                                     threads or os.cpu_count() or 1, grad)
        return grad

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_scene_free(self.h)
            self.h = None

    def intersect(self, rays, mode=TRAV_CULLED, stats=None):
        rays = _f32(rays)
        n = rays.shape[0]
        geo, tri, prim = (np.zeros(n, np.int32) for _ in range(3))
        t = np.zeros(n, np.float32)
        b = np.zeros((n, 2), np.float32)
        vtx = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        st = stats if stats is not None else None
        self.lib.orc_intersect(self.h, rays, n, mode, geo, tri, prim, t, b, vtx, nrm,
                               st.ctypes.data if st is not None else None)
        return dict(geo=geo, tri=tri, prim=prim, t=t, b=b, vertex=vtx, normal=nrm)

    def occluded(self, rays, tmin, tmax, mode=TRAV_REF, stats=None):
        rays = _f32(rays)
        n = rays.shape[0]
        out = np.zeros(n, np.uint8)
        self.lib.orc_occluded(self.h, rays, _f32(tmin), _f32(tmax), n, mode, out,
                              stats.ctypes.data if stats is not None else None)
        return out

    def bsdf_eval(self, mat, nrm, o, i, prec=0):
        n = nrm.shape[0]
        f = np.zeros((n, 3), np.float64)
        self.lib.orc_bsdf_eval(self.h, mat, n, _f32(nrm), _f32(o), _f32(i), prec, f)
        return f

    def bsdf_pdf(self, mat, nrm, o, i, prec=0):
        n = nrm.shape[0]
        p = np.zeros(n, np.float64)
        self.lib.orc_bsdf_pdf(self.h, mat, n, _f32(nrm), _f32(o), _f32(i), prec, p)
        return p

    def bsdf_sample(self, mat, seed, nrm, o, prec=0, ref_mixture=False):
        n = nrm.shape[0]
        i = np.zeros((n, 3), np.float64)
        d = np.zeros(n, np.float64)
        self.lib.orc_bsdf_sample(self.h, mat, seed, n, _f32(nrm), _f32(o), prec, int(ref_mixture), i, d)
        return i, d

    def bsdf_sample_u(self, mat, nrm, o, u, ref_mixture=True):
        """Sample k consumes u[k, :] in order (u: float32 [n, per_sample])."""
        n = nrm.shape[0]
        u = _f32(u)
        i = np.zeros((n, 3), np.float64)
        d = np.zeros(n, np.float64)
        self.lib.orc_bsdf_sample_u(self.h, mat, n, _f32(nrm), _f32(o), u, u.shape[1], int(ref_mixture), i, d)
        return i, d

    def pcg_draws(self, seed, pixel, sample, n):
        out = np.zeros(n, np.uint32)
        self.lib.orc_pcg_draws(seed, pixel, sample, n, out)
        return out

    def sincos_2pi(self, u):
        u = _f32(u)
        s = np.zeros_like(u)
        c = np.zeros_like(u)
        self.lib.orc_sincos_2pi(u.shape[0], u, s, c)
        return s, c

    def camera_rays(self, cam, w, h):
        rays = np.zeros((w * h, 6), np.float32)
        self.lib.orc_camera_rays(cam.sensor, cam.focal, cam.aspect, _f32(cam.transform), w, h, rays)
        return rays

    def render(self, tracer, width, height, spp, seed=1, max_depth=5, max_path_len=8, traversal=TRAV_CULLED,
               pixels: Optional[np.ndarray] = None, prec=0, threads=None, stats=None, camera=None):
        cam = (camera or self.desc.camera).packed()
        npix = width * height if pixels is None else int(pixels.shape[0])
        out = np.zeros((npix, 4), np.float32 if prec == 0 else np.float64)
        px = None if pixels is None else np.ascontiguousarray(pixels, np.uint32)
        threads = threads or os.cpu_count() or 1
        self.lib.orc_render(self.h, cam, tracer, max_depth, max_path_len, seed, traversal, width, height, spp,
                            px.ctypes.data if px is not None else None, npix, prec, threads, out.ctypes.data,
                            stats.ctypes.data if stats is not None else None)
        return out


class Reference(_Backend):
    prefix = "ref_"

    def __init__(self, desc=None):
        lib = C.CDLL(REF_SO)
        self.lib = lib
        lib.ref_scene_new.restype = C.c_void_p
        lib.ref_scene_free.argtypes = [C.c_void_p]
        lib.ref_add_oren_nayar.argtypes = [C.c_void_p, _f32p, C.c_float]
        lib.ref_add_cook_torr.argtypes = [C.c_void_p, _f32p, C.c_float, C.c_float, C.c_float]
        lib.ref_add_mixture.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float]
        lib.ref_add_mesh.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, _u32p, C.c_uint32, C.c_int]
        lib.ref_add_area_light.argtypes = [C.c_void_p, C.c_int, _f32p]
        lib.ref_add_sky_light.argtypes = [C.c_void_p, _f32p]
        lib.ref_light_info.argtypes = [C.c_void_p, C.c_int, _f32p]
        lib.ref_light_eval.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _f32p, _f32p, _f32p, _f32p]
        lib.ref_commit.argtypes = [C.c_void_p]
        lib.ref_bvh_size.argtypes = [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.ref_bvh_dump.argtypes = [C.c_void_p, _f32p, _u32p, _i32p]
        lib.ref_intersect.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p]
        lib.ref_occluded.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, C.c_uint32, C.c_int, _u8p]
        lib.ref_bsdf_eval.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _f32p, _f32p, _f32p, _f32p]
        lib.ref_bsdf_sample.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, _f32p, _f32p, _f32p, _f32p]
        lib.ref_rng_draws.argtypes = [C.c_uint32, C.c_uint32, _f32p]
        lib.ref_camera_rays.argtypes = [C.c_float, C.c_float, C.c_float, _f32p, C.c_uint32, C.c_uint32, _f32p]
        lib.ref_render.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float, _f32p, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        lib.ref_render.restype = C.c_uint32
        self.h = C.c_void_p(lib.ref_scene_new())
        self.desc = desc
        if desc is not None:
            self._load_desc(desc)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_scene_free(self.h)
            self.h = None

    def intersect(self, rays, linear=False):
        rays = _f32(rays)
        n = rays.shape[0]
        geo, tri = np.zeros(n, np.int32), np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        vtx = np.zeros((n, 3), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        self.lib.ref_intersect(self.h, rays, n, int(linear), geo, tri, t, vtx, nrm)
        return dict(geo=geo, tri=tri, t=t, vertex=vtx, normal=nrm)

    def occluded(self, rays, tmin, tmax, linear=False):
        rays = _f32(rays)
        out = np.zeros(rays.shape[0], np.uint8)
        self.lib.ref_occluded(self.h, rays, _f32(tmin), _f32(tmax), rays.shape[0], int(linear), out)
        return out

    def bsdf_eval(self, mat, nrm, o, i):
        n = nrm.shape[0]
        f = np.zeros((n, 3), np.float32)
        self.lib.ref_bsdf_eval(self.h, mat, n, _f32(nrm), _f32(o), _f32(i), f)
        return f

    def bsdf_sample(self, mat, seed, nrm, o):
        n = nrm.shape[0]
        i = np.zeros((n, 3), np.float32)
        d = np.zeros(n, np.float32)
        self.lib.ref_bsdf_sample(self.h, mat, seed, n, _f32(nrm), _f32(o), i, d)
        return i, d

    def rng_draws(self, seed, n):
        out = np.zeros(n, np.float32)
        self.lib.ref_rng_draws(seed, n, out)
        return out

    def camera_rays(self, cam, w, h):
        rays = np.zeros((w * h, 6), np.float32)
        self.lib.ref_camera_rays(cam.sensor, cam.focal, cam.aspect, _f32(cam.transform), w, h, rays)
        return rays

    def render(self, tracer, width, height, spp, threads=1, camera=None):
        cam = camera or self.desc.camera
        out = np.zeros((width * height, 4), np.float32)
        taken = self.lib.ref_render(self.h, tracer, cam.sensor, cam.focal, cam.aspect, _f32(cam.transform), width,
                                    height, spp, threads, out)
        return out, taken


_compose_lib = None


def _clib():
    global _compose_lib
    if _compose_lib is None:
        lib = C.CDLL(ORACLE_SO)
        args = [_f32p, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_float, _u8p, C.POINTER(C.c_float),
                C.POINTER(C.c_uint32)]
        lib.orc_compose.argtypes = args
        lib.orc_compose_libm.argtypes = args
        lib.orc_compose_math.argtypes = [C.c_int, _f32p, C.c_uint32, _f32p]
        _compose_lib = lib
    return _compose_lib


COMPOSE_CLAMP, COMPOSE_ACES = 0, 1


def compose(rgba, width, height, kind, auto_exposure=True, exposure=0.0, libm=False):
    """Film rgba [h*w, 4] float32 -> (frame uint8 [h*w, 4], exposure applied, pixels counted).
    libm=False: the library's arithmetic (own log / exp, fixed-order double sum); True: the
    reference's (std::pow / log / exp in float, float sum in pixel order)."""
    x = _f32(rgba).reshape(-1)
    assert x.size == 4 * width * height
    out = np.zeros((width * height, 4), np.uint8)
    e, n = C.c_float(), C.c_uint32()
    f = _clib().orc_compose_libm if libm else _clib().orc_compose
    f(x, width, height, kind, 1 if auto_exposure else 0, exposure, out.reshape(-1), C.byref(e), C.byref(n))
    return out, e.value, n.value


def compose_math(fn, x):
    """The compositor's own pow(x, 1/2.2f) (fn 0), log (1), exp (2) on float32 x."""
    x = _f32(x).reshape(-1)
    out = np.zeros_like(x)
    _clib().orc_compose_math(fn, x, x.size, out)
    return out


def ref_transform_mesh(transform, verts, normals):
    """The reference's trimesh::transform (oracle/_ref): (verts, normals) or None if it throws."""
    lib = C.CDLL(REF_SO)
    lib.ref_transform_mesh.argtypes = [_f32p, C.c_uint32, _f32p, _f32p, _f32p, _f32p]
    v, n = _f32(verts).reshape(-1, 3), _f32(normals).reshape(-1, 3)
    ov, on = np.zeros_like(v), np.zeros_like(n)
    rc = lib.ref_transform_mesh(_f32(transform).reshape(-1), v.shape[0], v.reshape(-1), n.reshape(-1),
                                ov.reshape(-1), on.reshape(-1))
    return None if rc != 0 else (ov, on)
