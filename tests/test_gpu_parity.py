"""GPU parity: the HIP library (through its C ABI) against the CPU oracle, bit for bit.

Closest hits (primitive, t, barycentrics, vertex, shading normal) must equal the oracle's —
which is pinned to the reference's own intersector in test_golden.py — in
both traversal modes; any-hit answers, BSDF eval / pdf / sample, camera rays, PCG streams and
whole films of every tracer must be identical as well.
"""

import numpy as np
import pytest

from x3dcore2_amd import scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


QUERY_SCENES = {
    "box": lambda: scenes.box_scene(),
    "soup": lambda: scenes.random_soup(4000, 3),
    "furnace": lambda: scenes.furnace_sphere(),
    "indirect_l3": lambda: scenes.indirect_scene(level=3),
    "spheres_small": lambda: scenes.spheres_scene(count=12, target_tris=30000, seed=77),
    "multi_light": lambda: scenes.multi_light_scene(),
    "sky": lambda: scenes.sky_scene(),
    "ties": lambda: scenes.tie_scene(),
}


@pytest.fixture(scope="module", params=sorted(QUERY_SCENES))
def scene_pair(request, x3d, oracle_mod):
    desc = QUERY_SCENES[request.param]()
    print(f"[gpu] scene {request.param}: {desc.num_tris} tris", flush=True)
    return desc, x3d.Scene(desc), oracle_mod.Oracle(desc)


def test_bvh_layout_matches_oracle(scene_pair):
    desc, gs, orc = scene_pair
    gb, gm, gp = gs.bvh()
    ob, om, op = orc.bvh()
    assert np.array_equal(_bits(gb), _bits(ob))
    assert np.array_equal(gm, om)
    assert np.array_equal(gp, op)


@pytest.mark.parametrize("traversal", [0, 1])
def test_closest_hit_bit_exact(scene_pair, x3d, oracle_mod, traversal):
    desc, gs, orc = scene_pair
    rays = scenes.seeded_rays(desc, 50000, 5)
    g = gs.intersect(rays, traversal)
    o = orc.intersect(rays, oracle_mod.TRAV_CULLED)
    assert (o["prim"] >= 0).mean() > 0.3
    assert np.array_equal(g["prim"], o["prim"])
    assert np.array_equal(g["geo"], o["geo"])
    assert np.array_equal(g["tri"], o["tri"])
    hit = o["prim"] >= 0
    assert np.array_equal(_bits(g["t"][hit]), _bits(o["t"][hit]))
    assert np.array_equal(_bits(np.stack([g["b1"], g["b2"]], 1)[hit]), _bits(o["b"][hit]))
    assert np.array_equal(_bits(g["vertex"][hit]), _bits(o["vertex"][hit]))
    assert np.array_equal(_bits(g["normal"][hit]), _bits(o["normal"][hit]))


def test_any_hit_matches(scene_pair, oracle_mod):
    desc, gs, orc = scene_pair
    rays = scenes.seeded_rays(desc, 30000, 9)
    o = orc.intersect(rays, oracle_mod.TRAV_CULLED)
    t = np.where(o["prim"] >= 0, o["t"], 2.0).astype(np.float32)
    rng = np.random.default_rng(2)
    scale = rng.choice([0.5, 0.999, 1.001, 2.0], size=t.shape[0]).astype(np.float32)
    tmin = np.full(t.shape, 1e-4, np.float32)
    tmax = (t * scale).astype(np.float32)
    g = gs.occluded(rays, tmin, tmax)
    ref = orc.occluded(rays, tmin, tmax)
    assert 0.2 < ref.mean() < 0.8
    assert np.array_equal(g, ref)


def _plane_rays(gs, n, seed):
    """Rays parallel to a coordinate plane whose origin lies exactly on a BVH box plane of the
    axis they do not move along: the slab test meets 0 * inf = NaN there, where the reference's
    std::max / std::min keep or drop the NaN by argument order."""
    rng = np.random.default_rng(seed)
    bounds, _, _ = gs.bvh()
    lo, hi = bounds[:, :3].min(0), bounds[:, 3:].max(0)
    node = rng.integers(0, bounds.shape[0], size=n)
    axis = rng.integers(0, 3, size=n)
    side = rng.integers(0, 2, size=n)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    o[np.arange(n), axis] = bounds[node, axis + 3 * side]
    d = rng.normal(size=(n, 3))
    d[np.arange(n), axis] = 0.0
    one = rng.random(n) < 0.5  # half of them move along a single axis
    keep = (axis + 1 + rng.integers(0, 2, size=n)) % 3
    d[one] = 0.0
    d[one, keep[one]] = rng.choice([-1.0, 1.0], size=one.sum())
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d.astype(np.float32)], axis=1).astype(np.float32)


@pytest.mark.parametrize("traversal", [0, 1])
def test_axis_parallel_rays_on_box_planes_bit_exact(scene_pair, oracle_mod, traversal):
    desc, gs, orc = scene_pair
    rays = _plane_rays(gs, 20000, 4)
    g = gs.intersect(rays, traversal)
    o = orc.intersect(rays, oracle_mod.TRAV_CULLED)
    assert (o["prim"] >= 0).mean() > 0.1
    assert np.array_equal(g["prim"], o["prim"])
    hit = o["prim"] >= 0
    assert np.array_equal(_bits(g["t"][hit]), _bits(o["t"][hit]))
    tmin = np.full(rays.shape[0], 1e-4, np.float32)
    tmax = np.where(hit, o["t"] * 1.001, 2.0).astype(np.float32)
    assert np.array_equal(gs.occluded(rays, tmin, tmax), orc.occluded(rays, tmin, tmax))


def _bsdf_inputs(n, seed):
    rng = np.random.default_rng(seed)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)

    def hemi():
        v = rng.normal(size=(n, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        flip = rng.random(n) < 0.9  # mostly upper hemisphere, some below
        return np.where(flip[:, None], v * np.sign((v * nrm).sum(1))[:, None], v)

    return nrm.astype(np.float32), hemi().astype(np.float32), hemi().astype(np.float32)


def _material_scene():
    desc = scenes.box_scene()
    base = len(desc.materials)
    desc.materials += [scenes.Material(scenes.COOK_TORR, (0.787, 0.787, 0.787), roughness=r, ior=scenes.GLOSSY_IOR)
                       for r in (0.05, 0.2, 0.4, 0.7, 0.8)]
    desc.materials += [scenes.Material(scenes.OREN_NAYAR, (0.725, 0.71, 0.68), sigma=0.078)]
    desc.materials += [scenes.Material(scenes.MIXTURE, sub0=0, sub1=base + 1, ratio=0.5)]
    return desc


@pytest.fixture(scope="module")
def mat_pair(x3d, oracle_mod):
    desc = _material_scene()
    return desc, x3d.Scene(desc), oracle_mod.Oracle(desc)


def test_bsdf_eval_pdf_sample_bit_exact(mat_pair):
    desc, gs, orc = mat_pair
    nrm, o, i = _bsdf_inputs(20000, 4)
    for m in [-1] + list(range(len(desc.materials))):
        print(f"[gpu] material {m}", flush=True)
        assert np.array_equal(_bits(gs.bsdf_eval(m, nrm, o, i)), _bits(orc.bsdf_eval(m, nrm, o, i)))
        assert np.array_equal(_bits(gs.bsdf_pdf(m, nrm, o, i)), _bits(orc.bsdf_pdf(m, nrm, o, i)))
        for refmix in (False, True):
            gi, gd = gs.bsdf_sample(m, 123, nrm, o, refmix)
            oi, od = orc.bsdf_sample(m, 123, nrm, o, ref_mixture=refmix)
            assert np.array_equal(_bits(gi), _bits(oi)), m
            assert np.array_equal(_bits(gd), _bits(od)), m


def test_camera_rays_and_pcg(x3d, oracle_mod):
    desc = scenes.box_scene()
    orc = oracle_mod.Oracle(desc)
    cam = desc.camera
    cam2 = scenes.Camera(0.032, 0.035, 4.0 / 3.0,
                         np.array([0.7071068, 0, -0.7071068, 0, 0, 1, 0, 0, 0.7071068, 0, 0.7071068, 0, 1, 2, 3, 1],
                                  np.float32))
    for c, (w, h) in ((cam, (64, 48)), (cam2, (1025, 769))):
        assert np.array_equal(_bits(x3d.camera_rays(c, w, h)), _bits(orc.camera_rays(c, w, h)))
    for seed, pix, s in ((1, 0, 0), (7, 12345, 63), (2 ** 40 + 3, 2 ** 33, 1000)):
        assert np.array_equal(x3d.pcg_draws(seed, pix, s, 256), orc.pcg_draws(seed, pix, s, 256))


RENDER_CASES = [
    ("box", "UNIDIRECT_MIS", dict(max_depth=5)),
    ("box", "UNIDIRECT_LT1", {}),
    ("box", "UNIDIRECT", {}),
    ("box", "DIRECT", {}),
    ("box", "NORMAL", {}),
    ("box", "POSITION", {}),
    ("spheres_small", "POSITION", {}),
    ("box", "BIDIRECT_LT2", {}),
    ("indirect_l3", "BIDIRECT_LT2", {}),
    ("spheres_small", "BIDIRECT_LT2", {}),
    ("furnace", "BIDIRECT_LT2", {}),
    ("spheres_small", "UNIDIRECT_MIS", dict(max_depth=8)),
    ("spheres_small", "UNIDIRECT_LT1", {}),
    ("box", "BIDIRECT_MIS", dict(max_path_len=8)),
    ("box", "BIDIRECT_MIS", dict(max_path_len=1)),
    ("indirect_l3", "BIDIRECT_MIS", dict(max_path_len=6)),
    ("spheres_small", "BIDIRECT_MIS", dict(max_path_len=5)),
    ("furnace", "BIDIRECT_MIS", dict(max_path_len=16)),
] + [
    # several lights of unequal power and area (one a fan of unequal triangles), and the sky light
    # beside an area light: light choice by power, triangle choice by area, the sky's sampling and
    # evaluation, its MIS weight, its emission for the light subpaths
    (name, tracer, kw)
    for name in ("multi_light", "sky")
    for tracer, kw in (("UNIDIRECT_MIS", dict(max_depth=6)), ("UNIDIRECT_LT1", {}), ("UNIDIRECT", {}),
                       ("DIRECT", {}), ("NORMAL", {}), ("POSITION", {}), ("BIDIRECT_LT2", {}),
                       ("BIDIRECT_MIS", dict(max_path_len=6)))
]


@pytest.mark.parametrize("scene_name,tracer,kw", RENDER_CASES)
def test_render_bit_exact_vs_oracle(x3d, oracle_mod, scene_name, tracer, kw):
    desc = QUERY_SCENES[scene_name]()
    print(f"[gpu] render {scene_name} {tracer}", flush=True)
    w, h, spp = 48, 40, 4
    pt = getattr(x3d, "PT_" + tracer)
    gs = x3d.Scene(desc)
    r = x3d.Renderer(pt, seed=11, **kw)
    img, st = r.render(gs, desc.camera, w, h, spp)
    orc = oracle_mod.Oracle(desc)
    ref = orc.render(pt, w, h, spp, seed=11, max_depth=kw.get("max_depth", 5), max_path_len=kw.get("max_path_len", 8))
    assert np.isfinite(img).all()
    assert img[:, :3].max() > 0
    assert np.array_equal(_bits(img), _bits(ref)), np.abs(img - ref).max()
    assert st.num_samples == w * h * spp


def test_render_deterministic_and_shape_independent(x3d):
    desc = scenes.box_scene()
    gs = x3d.Scene(desc)
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, seed=3)
    a, _ = r.render(gs, desc.camera, 96, 64, 8)
    b, _ = r.render(gs, desc.camera, 96, 64, 8)
    assert np.array_equal(_bits(a), _bits(b))


WAVEFRONT_CASES = [
    dict(),                                      # defaults: one generation, graph
    dict(workgroup_size=64, queue_capacity=3000),   # several generations, partial last one
    dict(workgroup_size=128, sort_key=1),           # material sort
    dict(workgroup_size=256, stack_depth=16, ray_order=1, queue_capacity=1000),  # rays sorted by key
]


@pytest.mark.parametrize("kw", WAVEFRONT_CASES)
@pytest.mark.parametrize("scene_name", ["box", "spheres_small", "multi_light", "sky"])
def test_wavefront_film_identical_to_plain(x3d, scene_name, kw):
    desc = QUERY_SCENES[scene_name]()
    print(f"[gpu] wavefront {scene_name} {kw}", flush=True)
    gs = x3d.Scene(desc)
    w, h, spp = 48, 40, 5
    plain, sp = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=6, seed=17).render(gs, desc.camera, w, h, spp)
    wf = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=6, seed=17, **kw)
    for _ in range(2):  # second call replays the cached graph
        img, st = wf.render(gs, desc.camera, w, h, spp)
        assert np.array_equal(_bits(img), _bits(plain)), np.abs(img - plain).max()
        assert st.closest_queries == sp.closest_queries
        assert st.anyhit_queries == sp.anyhit_queries


REF_WAVEFRONT_CASES = [
    dict(),                                                   # one generation, graph
    dict(workgroup_size=128, stack_depth=8, queue_capacity=3000),  # generations of 1 sample
    dict(profile=1),                                          # launched without a graph
]


@pytest.mark.parametrize("kw", REF_WAVEFRONT_CASES)
@pytest.mark.parametrize("tracer", ["UNIDIRECT_LT1", "UNIDIRECT"])
@pytest.mark.parametrize("scene_name", ["box", "spheres_small", "multi_light", "sky"])
def test_wavefront_reference_tracers_identical_to_plain(x3d, scene_name, tracer, kw):
    """The reference-mode lt1 and unidirect tracers (Russian roulette, up to 64 vertices) in the
    wavefront form: the same film bits and the same query counts as the plain kernels."""
    desc = QUERY_SCENES[scene_name]()
    print(f"[gpu] wavefront {tracer} {scene_name} {kw}", flush=True)
    gs = x3d.Scene(desc)
    pt = getattr(x3d, "PT_" + tracer)
    w, h, spp = 48, 40, 5
    plain, sp = x3d.Renderer(pt, seed=23).render(gs, desc.camera, w, h, spp)
    wf = x3d.Renderer(pt, form=x3d.FORM_WAVEFRONT, seed=23, **kw)
    for _ in range(2):  # second call replays the cached graph
        img, st = wf.render(gs, desc.camera, w, h, spp)
        assert np.array_equal(_bits(img), _bits(plain)), np.abs(img - plain).max()
        assert st.closest_queries == sp.closest_queries
        assert st.anyhit_queries == sp.anyhit_queries
    assert img[:, :3].max() > 0


BDPT_WAVEFRONT_CASES = [
    dict(max_path_len=6),                                          # one generation
    dict(max_path_len=1),                                          # emission ray only
    dict(max_path_len=2, queue_capacity=3000),                     # generations of 1 sample
    dict(max_path_len=8, workgroup_size=128, stack_depth=8),
    dict(max_path_len=16, profile=1),                              # deepest subpaths, timed launches
]


@pytest.mark.parametrize("kw", BDPT_WAVEFRONT_CASES)
@pytest.mark.parametrize("scene_name", ["box", "indirect_l3", "spheres_small", "multi_light", "sky", "furnace"])
def test_wavefront_bdpt_identical_to_plain(x3d, scene_name, kw):
    """The bidirectional tracer in the wavefront form (subpath vertices as structure-of-arrays
    rows, connections traced by the any-hit kernel): the plain kernel's film bits and query
    counts, in one or several generations."""
    desc = QUERY_SCENES[scene_name]()
    print(f"[gpu] wavefront BDPT {scene_name} {kw}", flush=True)
    gs = x3d.Scene(desc)
    w, h, spp = 48, 40, 4
    plain, sp = x3d.Renderer(x3d.PT_BIDIRECT_MIS, seed=29, max_path_len=kw["max_path_len"]).render(
        gs, desc.camera, w, h, spp)
    wf = x3d.Renderer(x3d.PT_BIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, seed=29, **kw)
    for _ in range(2):
        img, st = wf.render(gs, desc.camera, w, h, spp)
        assert np.array_equal(_bits(img), _bits(plain)), np.abs(img - plain).max()
        assert st.closest_queries == sp.closest_queries
        assert st.anyhit_queries == sp.anyhit_queries
    assert img[:, :3].max() > 0
    if kw.get("profile"):
        assert st.trace_launches == 2 * kw["max_path_len"] and st.extend_ms > 0 and st.shadow_ms > 0


@pytest.mark.parametrize("tracer,small,large", [
    # the box room (34 triangles): from 2^24 samples on (MIS, lt1, unidirect), 2^22 (BDPT)
    ("UNIDIRECT_MIS", (48, 40, 5), (1024, 1024, 16)),
    ("UNIDIRECT_LT1", (48, 40, 5), (1024, 1024, 16)),
    ("UNIDIRECT", (48, 40, 5), (1024, 1024, 16)),
    ("BIDIRECT_MIS", (48, 40, 5), (1024, 1024, 4)),
])
def test_auto_form_picks_by_size_with_the_same_film(x3d, tracer, small, large):
    """X3D_FORM_AUTO renders small films in the plain form and large ones in the wavefront form,
    and either way the film equals the plain form's bit for bit."""
    desc = QUERY_SCENES["box"]()
    gs = x3d.Scene(desc)
    pt = getattr(x3d, "PT_" + tracer)
    kw = dict(max_path_len=6) if tracer == "BIDIRECT_MIS" else {}
    auto = x3d.Renderer(pt, form=x3d.FORM_AUTO, seed=31, **kw)
    plain = x3d.Renderer(pt, form=x3d.FORM_PLAIN, seed=31, **kw)
    for (w, h, spp), want in ((small, x3d.FORM_PLAIN), (large, x3d.FORM_WAVEFRONT)):
        print(f"[gpu] auto form {tracer} {w}x{h}x{spp}", flush=True)
        a, sa = auto.render(gs, desc.camera, w, h, spp)
        p, sp = plain.render(gs, desc.camera, w, h, spp)
        assert sa.form == want and sp.form == x3d.FORM_PLAIN
        assert np.array_equal(_bits(a), _bits(p))


@pytest.mark.parametrize("tracer", ["UNIDIRECT_LT1", "UNIDIRECT", "UNIDIRECT_MIS", "BIDIRECT_MIS"])
def test_wavefront_without_lights_and_ragged(x3d, tracer):
    """Edge cases of the wavefront form: a scene without lights (no light sample is drawn, the film
    is black) and a ragged film (7 x 3 pixels, 3 samples, one sample per generation) match the plain
    kernel bit for bit."""
    import dataclasses

    desc = QUERY_SCENES["box"]()
    pt = getattr(x3d, "PT_" + tracer)
    dark = dataclasses.replace(desc, lights=[])
    for d, (w, h, spp, cap) in ((dark, (24, 16, 2, 0)), (desc, (7, 3, 3, 1))):
        print(f"[gpu] wavefront edge {tracer} lights={len(d.lights)} {w}x{h}x{spp} cap={cap}", flush=True)
        gs = x3d.Scene(d)
        plain, sp = x3d.Renderer(pt, seed=9).render(gs, d.camera, w, h, spp)
        img, st = x3d.Renderer(pt, form=x3d.FORM_WAVEFRONT, seed=9, queue_capacity=cap).render(gs, d.camera, w, h, spp)
        assert np.array_equal(_bits(img), _bits(plain)), np.abs(img - plain).max()
        assert st.closest_queries == sp.closest_queries and st.anyhit_queries == sp.anyhit_queries
        if not d.lights:
            assert not img[:, :3].any() and st.anyhit_queries == 0


def test_wavefront_lt1_autotune(x3d):
    """The autotuner of an lt1 wavefront renderer tries the unsorted configurations only."""
    desc = QUERY_SCENES["spheres_small"]()
    gs = x3d.Scene(desc)
    w, h, spp = 64, 48, 3
    plain, _ = x3d.Renderer(x3d.PT_UNIDIRECT_LT1, seed=5).render(gs, desc.camera, w, h, spp)
    wf = x3d.Renderer(x3d.PT_UNIDIRECT_LT1, form=x3d.FORM_WAVEFRONT, seed=5)
    rep = wf.autotune(gs, desc.camera, w, h, 1)
    assert rep.candidates == 6 and rep.sort_key == x3d.SORT_NONE
    img, _ = wf.render(gs, desc.camera, w, h, spp)
    assert np.array_equal(_bits(img), _bits(plain))


def test_wavefront_autotune_and_profile(x3d):
    desc = QUERY_SCENES["spheres_small"]()
    gs = x3d.Scene(desc)
    w, h, spp = 64, 48, 3
    plain, _ = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=8, seed=5).render(gs, desc.camera, w, h, spp)
    wf = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=8, seed=5)
    rep = wf.autotune(gs, desc.camera, w, h, 1)
    # 3072 probe paths: stage 2 (capacities of 2M paths and more) is skipped
    assert rep.candidates == 12 and rep.queue_capacity == 1 << 26 and rep.best_ms > 0 and rep.best_ms <= rep.default_ms * 1.5
    assert rep.workgroup_size in (64, 128, 256) and rep.stack_depth in (8, 16)
    img, _ = wf.render(gs, desc.camera, w, h, spp)
    assert np.array_equal(_bits(img), _bits(plain))
    prof = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=8, seed=5, profile=1,
                        stack_depth=8, workgroup_size=64)
    img, st = prof.render(gs, desc.camera, w, h, spp)
    assert np.array_equal(_bits(img), _bits(plain))
    assert st.trace_launches == 8 and st.extend_ms > 0 and st.shadow_ms > 0


def test_bdpt_generations_and_edge_cases(x3d, oracle_mod):
    """Bidirectional tracer split into several generations of samples (5 spp in generations of
    2, 2, 1 samples) gives the oracle's film; a scene without lights renders black."""
    desc = QUERY_SCENES["box"]()
    w, h, spp = 48, 40, 5
    gs = x3d.Scene(desc)
    orc = oracle_mod.Oracle(desc)
    ref = orc.render(x3d.PT_BIDIRECT_MIS, w, h, spp, seed=4, max_path_len=7)
    for cap in (w * h * 2, w * h * 2 + 17, 1):
        print(f"[gpu] bdpt generations, capacity {cap}", flush=True)
        img, st = x3d.Renderer(x3d.PT_BIDIRECT_MIS, max_path_len=7, seed=4, queue_capacity=cap).render(
            gs, desc.camera, w, h, spp)
        assert np.array_equal(_bits(img), _bits(ref)), np.abs(img - ref).max()
        assert st.num_samples == w * h * spp
    dark = scenes.box_scene()
    dark.lights = []
    img, _ = x3d.Renderer(x3d.PT_BIDIRECT_MIS, seed=4).render(x3d.Scene(dark), dark.camera, 16, 16, 2)
    assert not img[:, :3].any()


def test_plain_mis_both_work_shapes_bit_exact(x3d, oracle_mod):
    """The plain MIS tracer runs over (pixel, sample) items of generations sized by the queue
    capacity: a film with more pixels than the chip has lanes and a small film with more samples
    per pixel both give the oracle's film."""
    desc = QUERY_SCENES["spheres_small"]()
    gs = x3d.Scene(desc)
    orc = oracle_mod.Oracle(desc)
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=6, seed=21)
    for (w, h, spp) in ((1280, 1024, 2), (200, 150, 6)):
        print(f"[gpu] plain MIS {w}x{h}x{spp}", flush=True)
        img, _ = r.render(gs, desc.camera, w, h, spp)
        pix = np.random.default_rng(w).choice(w * h, size=3000, replace=False).astype(np.uint32)
        ref = orc.render(x3d.PT_UNIDIRECT_MIS, w, h, spp, seed=21, max_depth=6, pixels=pix)
        assert np.array_equal(_bits(img[pix]), _bits(ref))
