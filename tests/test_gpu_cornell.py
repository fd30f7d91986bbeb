"""The reference's own Cornell box (tst_renderer.cpp:37-115) loaded through x3d_scene_add_obj from
its OBJ files, with its materials (Oren-Nayar sigma = 0.078, cook_torr r = 0.2 with
eta = 2.93 + 3i) and light:

- the BVH and closest / any hits equal the fixtures written from the reference's own build;
- tst_renderer's check (800 x 600, 1 spp, 10 renders, unidirect_lt1): every pixel finite and
  non-negative, irradiance above 5000 per channel;
- films bit-identical to the oracle's (unidirect_lt1 and the MIS tracer);
- unidirect_lt1 against the reference renderer's own, within Monte Carlo error.
"""

import os

import numpy as np
import pytest

from mc_stats import gpu_replicates, split_z
from x3dcore2_amd import scenes

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PARTS = ["left_wall", "right_wall", "back_wall", "ceiling", "floor", "left_sphere", "right_sphere", "light"]
MAT_OF = [1, 2, 0, 0, 0, 3, 0, 4]


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _materials():
    return [scenes.Material(scenes.OREN_NAYAR, (0.725, 0.710, 0.680), sigma=0.078),
            scenes.Material(scenes.OREN_NAYAR, (0.630, 0.065, 0.050), sigma=0.078),
            scenes.Material(scenes.OREN_NAYAR, (0.140, 0.450, 0.091), sigma=0.078),
            scenes.Material(scenes.COOK_TORR, (0.95, 0.95, 0.95), roughness=0.2, ior=(2.93, 3.0)),
            scenes.Material(scenes.OREN_NAYAR, (0.0, 0.0, 0.0), sigma=0.078)]


def _camera():
    m = np.eye(4, dtype=np.float32)
    m[0, 3], m[1, 3], m[2, 3] = 0.0, 0.795, 3.4
    return scenes.Camera(0.032, 0.035, 4.0 / 3.0, m.T.reshape(-1).copy())


RADIANCE = tuple(float(np.float32(c) * np.float32(15.0)) for c in (0.911, 0.660, 0.345))


@pytest.fixture(scope="module")
def cornell(x3d):
    sc = x3d.Scene()
    ids = [sc.add_material(m) for m in _materials()]
    mesh_ids = [sc.add_obj(os.path.join(GOLD, "cornellbox", p + ".obj"), ids[MAT_OF[k]]) for k, p in enumerate(PARTS)]
    sc.add_area_light(mesh_ids[-1], RADIANCE)
    sc.commit()
    meshes = []
    for k, p in enumerate(PARTS):
        v, n, _, t = x3d.load_obj(os.path.join(GOLD, "cornellbox", p + ".obj"))
        meshes.append(scenes.Mesh(v, n, t, MAT_OF[k]))
    desc = scenes.SceneDesc("cornell", _materials(), meshes, [(7, RADIANCE)], _camera(), 800, 600, 1, 5)
    print(f"[gpu] cornell box from OBJ: {desc.num_tris} tris", flush=True)
    return sc, desc


def test_bvh_and_hits_match_reference_fixtures(cornell):
    sc, desc = cornell
    z = np.load(os.path.join(GOLD, "scene_cornell.npz"))
    b, m, p = sc.bvh()
    assert np.array_equal(_bits(b), _bits(z["bounds"]))
    assert np.array_equal(m, z["meta"]) and np.array_equal(p, z["prims"])
    h = sc.intersect(z["rays"])
    assert np.array_equal(h["geo"], z["hit_geo"]) and np.array_equal(h["tri"], z["hit_tri"])
    hit = z["hit_geo"] >= 0
    assert np.array_equal(_bits(h["t"][hit]), _bits(z["hit_t"][hit]))
    assert np.array_equal(_bits(h["vertex"][hit]), _bits(z["hit_vertex"][hit]))
    assert np.array_equal(_bits(h["normal"][hit]), _bits(z["hit_normal"][hit]))
    occ = sc.occluded(z["rays"], z["occ_tmin"], z["occ_tmax"])
    assert np.array_equal(occ, z["occ"])


def test_tst_renderer_check(x3d, cornell):
    """pt_render_cornel_balls (tst_renderer.cpp:117-167): 10 renders of 800 x 600 at 1 spp."""
    sc, desc = cornell
    for k in range(10):
        r = x3d.Renderer(x3d.PT_UNIDIRECT_LT1, seed=k)
        img, _ = r.render(sc, desc.camera, 800, 600, 1)
        rgb = img[:, :3]
        assert np.isfinite(rgb).all() and (rgb >= 0).all(), k
        irr = rgb.astype(np.float64).sum(0)
        assert (irr > 5000.0).all(), (k, irr)
    print(f"[gpu] cornell 800x600x1 irradiance {irr}", flush=True)


@pytest.mark.parametrize("tracer", ["lt1", "mis"])
def test_films_bit_identical_to_oracle(x3d, oracle_mod, cornell, tracer):
    sc, desc = cornell
    w, h, spp = 96, 72, 4
    if tracer == "lt1":
        img, _ = x3d.Renderer(x3d.PT_UNIDIRECT_LT1, seed=3).render(sc, desc.camera, w, h, spp)
        ref = oracle_mod.Oracle(desc).render(oracle_mod.TR_UNIDIRECT_LT1, w, h, spp, seed=3)
    else:
        img, _ = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=5, seed=3).render(sc, desc.camera, w, h, spp)
        ref = oracle_mod.Oracle(desc).render(oracle_mod.TR_UNIDIRECT_MIS, w, h, spp, seed=3, max_depth=5)
    assert np.array_equal(_bits(img), _bits(ref)), np.abs(img - ref).max()


def test_lt1_vs_reference_renderer(x3d, ref_mod, cornell):
    """GPU unidirect_lt1 (PCG streams) against the reference renderer's own unidirect_lt1 on its
    Cornell box, per pixel, within the combined Monte Carlo error."""
    sc, desc = cornell
    w, h, spp, reps = 48, 36, 64, 32
    imgs = gpu_replicates(x3d, desc, x3d.PT_UNIDIRECT_LT1, w, h, spp, reps, scene=sc)
    ref = ref_mod.Reference(desc)
    tasks = 64
    img, taken = ref.render(ref_mod.TR_UNIDIRECT_LT1, w, h, 4 * tasks, threads=tasks)
    assert taken == 4 * tasks
    z, rel_mse = split_z(imgs, spp, img[:, :3].astype(np.float64), taken)
    print(f"[gpu] cornell lt1 vs reference renderer: rel MSE {rel_mse:.2e}, |z|>3 in {np.mean(np.abs(z) > 3):.4f}, "
          f"mean z {z.mean():.3f}", flush=True)
    assert np.mean(np.abs(z) > 3) < 0.03
    assert abs(z.mean()) < 4 / np.sqrt(z.size) + 0.05
