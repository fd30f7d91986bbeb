"""Films with the firefly filter (x3d_render's firefly_filter, sampling_task::run,
renderer.cpp:46-83) bit-identical to the oracle's (orc_render_firefly: every sample image
rendered, filtered, summed per pixel in sample order), for every tracer and both forms of the
MIS, lt1, unidirect and bidirectional tracers, in one generation and in several. The filter only lowers samples, so the filtered
film is at most the unfiltered one everywhere and below it on the pixels it touched; the scenes
and sizes are chosen so the oracle caps samples on dozens of pixels.
"""

import numpy as np
import pytest

from x3dcore2_amd import scenes

pytestmark = pytest.mark.gpu

SCENES = {
    "indirect_l3": lambda: scenes.indirect_scene(level=3),
    "spheres_small": lambda: scenes.spheres_scene(count=12, target_tris=30000, seed=77),
}


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


CASES = [
    ("indirect_l3", "UNIDIRECT_MIS", dict()),
    ("indirect_l3", "UNIDIRECT_MIS", dict(queue_capacity=3000)),            # generations of 1 sample
    ("indirect_l3", "UNIDIRECT_MIS", dict(form="wavefront")),
    ("indirect_l3", "UNIDIRECT_MIS", dict(form="wavefront", queue_capacity=5000)),
    ("spheres_small", "UNIDIRECT_MIS", dict(form="wavefront", max_depth=8)),
    ("indirect_l3", "UNIDIRECT_LT1", dict()),
    ("spheres_small", "UNIDIRECT_LT1", dict(queue_capacity=2000)),
    ("indirect_l3", "UNIDIRECT", dict()),
    ("indirect_l3", "UNIDIRECT_LT1", dict(form="wavefront")),
    ("spheres_small", "UNIDIRECT", dict(form="wavefront", queue_capacity=3000)),
    ("indirect_l3", "DIRECT", dict()),
    ("indirect_l3", "NORMAL", dict()),
    ("indirect_l3", "BIDIRECT_LT2", dict()),
    ("indirect_l3", "BIDIRECT_MIS", dict(max_path_len=6)),
    ("spheres_small", "BIDIRECT_MIS", dict(max_path_len=5, queue_capacity=4000)),
    ("indirect_l3", "BIDIRECT_MIS", dict(form="wavefront", max_path_len=6)),
    ("spheres_small", "BIDIRECT_MIS", dict(form="wavefront", max_path_len=5, queue_capacity=4000)),
]


@pytest.mark.parametrize("scene_name,tracer,kw", CASES)
def test_firefly_film_bit_exact_vs_oracle(x3d, oracle_mod, scene_name, tracer, kw):
    desc = SCENES[scene_name]()
    print(f"[gpu] firefly {scene_name} {tracer} {kw}", flush=True)
    w, h, spp = 48, 40, 4
    kw = dict(kw)
    if kw.pop("form", None) == "wavefront":
        kw["form"] = x3d.FORM_WAVEFRONT
    pt = getattr(x3d, "PT_" + tracer)
    gs = x3d.Scene(desc)
    r = x3d.Renderer(pt, seed=11, **kw)
    img, st = r.render(gs, desc.camera, w, h, spp, firefly_filter=True)
    plain, _ = r.render(gs, desc.camera, w, h, spp)
    orc = oracle_mod.Oracle(desc)
    md, mpl = kw.get("max_depth", 5), kw.get("max_path_len", 8)
    ref = orc.render_firefly(pt, w, h, spp, seed=11, max_depth=md, max_path_len=mpl)
    ref_plain = orc.render(pt, w, h, spp, seed=11, max_depth=md, max_path_len=mpl)
    assert np.array_equal(_bits(img), _bits(ref)), np.abs(img - ref).max()
    assert np.array_equal(_bits(plain), _bits(ref_plain))
    assert st.num_samples == w * h * spp
    assert (img <= plain).all()
    touched = (img[:, :3] != plain[:, :3]).any(1).sum()
    print(f"[gpu]   {touched} pixels filtered", flush=True)
    if tracer != "NORMAL":   # normal colours stay in [0, 1]: nothing to cap
        assert touched > 0
    else:
        assert touched == 0
