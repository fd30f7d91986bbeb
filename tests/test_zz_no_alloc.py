"""x3d_render_device_noalloc (working memory from the library's fixed device array, no
allocation in the call) against x3d_render_device: identical films through the C ABI on two
calls in a row, the second with a smaller film, for the plain MIS, BDPT and lt1 tracers and a
renderer of the wavefront form; a film too large for the fixed array is refused."""

import ctypes as C

import numpy as np
import pytest

from x3dcore2_amd import scenes
from x3dcore2_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _noalloc(x3d, r, scene, cam, w, h, spp, d_ptr):
    st = x3d.RenderStats()
    rc = x3d.lib().x3d_render_device_noalloc(r.h, scene.h, C.byref(x3d.camera_desc(cam)), w, h, spp,
                                             C.c_void_p(d_ptr), C.byref(st))
    return rc, st


@pytest.mark.parametrize("scene_name,tracer,kw", [
    ("spheres_small", "UNIDIRECT_MIS", dict(max_depth=6)),
    ("box", "BIDIRECT_MIS", dict(max_path_len=6)),
    ("multi_light", "UNIDIRECT_LT1", {}),
    ("box", "UNIDIRECT_MIS", dict(max_depth=5, form=1)),
])
def test_noalloc_equals_render_device(x3d, scene_name, tracer, kw):
    desc = {"box": scenes.box_scene, "multi_light": scenes.multi_light_scene,
            "spheres_small": lambda: scenes.spheres_scene(count=12, target_tris=30000, seed=77)}[scene_name]()
    gs = x3d.Scene(desc)
    pt = getattr(x3d, "PT_" + tracer)
    r = x3d.Renderer(pt, seed=13, **kw)
    for (w, h, spp) in ((640, 480, 5), (97, 33, 3)):  # second call smaller
        print(f"[gpu] noalloc {scene_name} {tracer} {w}x{h}x{spp}", flush=True)
        a_d, b_d = DeviceArray(w * h * 16), DeviceArray(w * h * 16)
        r.render_device(gs, desc.camera, w, h, spp, a_d.ptr)
        rc, st = _noalloc(x3d, r, gs, desc.camera, w, h, spp, b_d.ptr)
        assert rc == 0, x3d.lib().x3d_last_error().decode()
        a = a_d.to_host(np.float32, (w * h, 4))
        b = b_d.to_host(np.float32, (w * h, 4))
        assert np.array_equal(_bits(a), _bits(b)), np.abs(a - b).max()
        assert b[:, :3].max() > 0 and st.num_samples == w * h * spp


def test_noalloc_full_size_spheres_film(x3d):
    """1024 x 1024 on the 1.3M-triangle scene fits (fewer resident workgroups for the stack spill)."""
    desc = scenes.spheres_scene()
    gs = x3d.Scene(desc)
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=8, seed=2)
    w, h, spp = 1024, 1024, 2
    a_d, b_d = DeviceArray(w * h * 16), DeviceArray(w * h * 16)
    r.render_device(gs, desc.camera, w, h, spp, a_d.ptr)
    rc, st = _noalloc(x3d, r, gs, desc.camera, w, h, spp, b_d.ptr)
    assert rc == 0, x3d.lib().x3d_last_error().decode()
    print(f"[gpu] noalloc full spheres: {st.kernel_ms:.1f} ms", flush=True)
    assert np.array_equal(_bits(a_d.to_host(np.float32, (w * h, 4))), _bits(b_d.to_host(np.float32, (w * h, 4))))


def test_noalloc_refuses_too_large_film(x3d):
    desc = scenes.box_scene()
    gs = x3d.Scene(desc)
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, seed=1)
    d = DeviceArray(16)
    rc, _ = _noalloc(x3d, r, gs, desc.camera, 2048, 1024, 1, d.ptr)
    assert rc == 1  # X3D_ERR_INVALID_ARGUMENT
    assert "fixed scratch memory" in x3d.lib().x3d_last_error().decode()
