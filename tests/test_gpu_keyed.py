"""Objects by id on the GPU: the if_obj_actuator protocol (load / unload / commit, obj.h:46-55)
as objdb::push_updates drives it (objdb.cpp:59-87), against scenes built by index:

- objects loaded in shuffled order under scattered ids commit in id order (the actuators'
  std::map order, pathspace.cpp:292-309, lightsources.cpp:45-52): BVH, hits and films equal an
  indexed scene built in that order, and x3d_scene_geometry_obj_id maps hits back to ids;
- a second load of a loaded id changes nothing (map insert);
- edits and a second commit: an unloaded geometry disappears, a geometry whose material was
  unloaded takes the fail-safe material, a reloaded geometry with a transform equals its
  x3d_transform_mesh copy; renderers created before the edit (plain and wavefront, whose
  graph was captured on the old scene) give the new scene's film;
- the position tracer keeps the path space's grown bounds after an unload (if_path_space::load
  never shrinks them).
"""

import numpy as np
import pytest

from x3dcore2_amd import scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _col_major(m):
    return np.ascontiguousarray(np.asarray(m, np.float32).T).reshape(-1)


def _indexed(x3d, mats, meshes, lights):
    """mats: [Material]; meshes: [(verts, normals, tris, material index or -1)]; lights:
    [(mesh index, radiance)] -- committed in list order."""
    s = x3d.Scene()
    for m in mats:
        s.add_material(m)
    for v, n, t, mi in meshes:
        s.add_mesh(v, n, t, mi)
    for mi, rad in lights:
        s.add_area_light(mi, rad)
    s.commit()
    return s


def _films(x3d, scene, cam, renderers=None):
    w, h, spp = 48, 40, 3
    rs = renderers or [x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=5, seed=7),
                       x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=5, seed=7)]
    return [r.render(scene, cam, w, h, spp)[0] for r in rs], rs


def _same_scene(a, b, desc):
    ba, ma, pa = a.bvh()
    bb, mb, pb = b.bvh()
    assert np.array_equal(_bits(ba), _bits(bb)) and np.array_equal(ma, mb) and np.array_equal(pa, pb)
    rays = scenes.seeded_rays(desc, 20000, 3)
    ha, hb = a.intersect(rays), b.intersect(rays)
    assert np.array_equal(ha["geo"], hb["geo"]) and np.array_equal(ha["tri"], hb["tri"])
    assert np.array_equal(_bits(ha["t"]), _bits(hb["t"]))


@pytest.fixture(scope="module")
def setup(x3d):
    desc = scenes.indirect_scene(level=2)
    rng = np.random.default_rng(12)
    mat_ids = [1000 - 7 * k for k in range(len(desc.materials))]      # id order reverses index order
    geo_ids = list(rng.permutation(len(desc.meshes)) * 5 + 40)
    light_ids = [900 + k for k in range(len(desc.lights))]
    keyed = x3d.Scene()
    ops = [("m", k) for k in range(len(desc.materials))] + [("g", k) for k in range(len(desc.meshes))] + \
          [("l", k) for k in range(len(desc.lights))]
    for kind, k in [ops[i] for i in rng.permutation(len(ops))]:
        if kind == "m":
            keyed.load_material(mat_ids[k], desc.materials[k])
        elif kind == "g":
            m = desc.meshes[k]
            keyed.load_geometry(geo_ids[k], m.verts, m.normals, m.tris,
                                material=mat_ids[m.material] if m.material >= 0 else None)
        else:
            mi, rad = desc.lights[k]
            keyed.load_light(light_ids[k], geo_ids[mi], rad)
    keyed.commit()
    return desc, keyed, mat_ids, geo_ids, light_ids


def _reference_order(desc, mat_ids, geo_ids, light_ids, mesh_override=None, drop_geo=(), drop_mat=()):
    mats = sorted((i, k) for k, i in enumerate(mat_ids) if i not in drop_mat)
    mat_rank = {k: r for r, (_, k) in enumerate(mats)}
    geos = sorted((i, k) for k, i in enumerate(geo_ids) if i not in drop_geo)
    geo_rank = {k: r for r, (_, k) in enumerate(geos)}
    meshes = []
    for _, k in geos:
        m = desc.meshes[k]
        v, n = (m.verts, m.normals) if mesh_override is None or k not in mesh_override else mesh_override[k]
        meshes.append((v, n, m.tris, mat_rank.get(m.material, -1) if m.material >= 0 else -1))
    lights = [(geo_rank[desc.lights[k][0]], desc.lights[k][1]) for _, k in sorted(zip(light_ids, range(len(light_ids))))]
    return [desc.materials[k] for _, k in mats], meshes, lights, [i for i, _ in geos]


def test_keyed_commits_in_id_order(x3d, setup):
    desc, keyed, mat_ids, geo_ids, light_ids = setup
    mats, meshes, lights, order = _reference_order(desc, mat_ids, geo_ids, light_ids)
    ref = _indexed(x3d, mats, meshes, lights)
    _same_scene(keyed, ref, desc)
    fk, _ = _films(x3d, keyed, desc.camera)
    fr, _ = _films(x3d, ref, desc.camera)
    for a, b in zip(fk, fr):
        assert np.array_equal(_bits(a), _bits(b))
    assert [keyed.geometry_obj_id(g) for g in range(len(order))] == order


def test_second_load_of_an_id_changes_nothing(x3d, setup):
    desc, keyed, mat_ids, geo_ids, light_ids = setup
    b0 = keyed.bvh()
    m = desc.meshes[0]
    keyed.load_geometry(geo_ids[0], m.verts * 2.0, m.normals, m.tris)
    keyed.load_material(mat_ids[0], scenes.Material(scenes.OREN_NAYAR, (0.1, 0.1, 0.1)))
    keyed.commit()
    b1 = keyed.bvh()
    assert all(np.array_equal(_bits(x), _bits(y)) for x, y in zip(b0, b1))


def test_edit_and_recommit(x3d, setup):
    desc, keyed, mat_ids, geo_ids, light_ids = setup
    films0, rs = _films(x3d, keyed, desc.camera)
    # unload a sphere and the material of another mesh; move a third mesh with a transform
    gone = geo_ids[8]
    mat_gone = mat_ids[desc.meshes[7].material]
    moved = 9
    t = np.eye(4, dtype=np.float32)
    t[:3, :3] = [[0.8, 0.0, 0.1], [0.0, 1.1, 0.0], [-0.1, 0.0, 0.9]]
    t[:3, 3] = [0.05, -0.02, 0.1]
    keyed.unload(x3d.OBJ_GEOMETRY, gone)
    keyed.unload(x3d.OBJ_MATERIAL, mat_gone)
    keyed.unload(x3d.OBJ_GEOMETRY, geo_ids[moved])
    m = desc.meshes[moved]
    keyed.load_geometry(geo_ids[moved], m.verts, m.normals, m.tris, material=mat_ids[m.material],
                        transform=_col_major(t))
    keyed.commit()
    tv, tn = x3d.transform_mesh(_col_major(t), m.verts, m.normals)
    mats, meshes, lights, order = _reference_order(desc, mat_ids, geo_ids, light_ids, mesh_override={moved: (tv, tn)},
                                                   drop_geo=(gone,), drop_mat=(mat_gone,))
    ref = _indexed(x3d, mats, meshes, lights)
    _same_scene(keyed, ref, desc)
    films1, _ = _films(x3d, keyed, desc.camera, rs)   # renderers and graphs from before the edit
    films_ref, _ = _films(x3d, ref, desc.camera)
    for a, b, c in zip(films1, films_ref, films0):
        assert np.array_equal(_bits(a), _bits(b))
        assert not np.array_equal(_bits(a), _bits(c))


def test_position_bounds_only_grow(x3d):
    desc = scenes.box_scene()
    s = x3d.Scene()
    all_verts = []
    for k, m in enumerate(desc.meshes):
        s.load_geometry(k, m.verts, m.normals, m.tris)
        all_verts.append(m.verts)
    big = np.array([[-5, -5, -5], [6, 7, 8], [0, 0, 9]], np.float32)
    s.load_geometry(99, big, np.tile([0, 0, 1], (3, 1)).astype(np.float32), np.array([[0, 1, 2]], np.uint32))
    all_verts.append(big)
    s.commit()
    s.unload(x3d.OBJ_GEOMETRY, 99)
    s.commit()
    w, h = 40, 32
    img, _ = x3d.Renderer(x3d.PT_POSITION).render(s, desc.camera, w, h, 1)
    rays = x3d.camera_rays(desc.camera, w, h).reshape(-1, 6)
    hits = s.intersect(rays)
    v = np.concatenate(all_verts)
    lo, hi = v.min(0), v.max(0)
    rng_ = (hi - lo).astype(np.float32)
    ok = hits["geo"] >= 0
    want = ((hits["vertex"][ok] - lo) / rng_).astype(np.float32)
    got = img[ok, :3]
    # the first-hit cache drops back-facing hits (compute_first_hit): compare the rest
    front = got.any(1)
    assert front.mean() > 0.8
    assert np.array_equal(_bits(got[front]), _bits(want[front]))


def test_material_params_follow_ids_after_edits(x3d):
    """x3d_scene_set_material_params on a keyed scene writes back by obj id: after
    load -> commit -> load_material(new lower id) -> set_material_params -> commit, the changed
    material is the one the caller addressed (committed index -> id of the last commit), and the
    film equals an indexed scene with the same parameters."""
    desc = scenes.box_scene()
    keyed = x3d.Scene()
    ids = [500 + 10 * k for k in range(len(desc.materials))]
    for i, m in zip(ids, desc.materials):
        keyed.load_material(i, m)
    for k, m in enumerate(desc.meshes):
        keyed.load_geometry(100 + k, m.verts, m.normals, m.tris, material=ids[m.material])
    for k, (mi, rad) in enumerate(desc.lights):
        keyed.load_light(900 + k, 100 + mi, rad)
    keyed.commit()
    # a new material under a lower id than all others: its rank is 0 at the next commit
    keyed.load_material(1, scenes.Material(scenes.OREN_NAYAR, (0.2, 0.9, 0.2)))
    new_albedo = np.array([[0.9, 0.1, 0.1]], np.float32)
    keyed.set_material_params(1, new_albedo)  # committed index 1 = obj id ids[1] (the red wall)
    keyed.commit()
    mats = [scenes.Material(scenes.OREN_NAYAR, (0.2, 0.9, 0.2))] + list(desc.materials)
    mats[2] = scenes.Material(scenes.OREN_NAYAR, tuple(float(a) for a in new_albedo[0]))
    indexed = _indexed(x3d, mats, [(m.verts, m.normals, m.tris, m.material + 1) for m in desc.meshes],
                       desc.lights)
    a, _ = _films(x3d, keyed, desc.camera)
    b, _ = _films(x3d, indexed, desc.camera)
    for fa, fb in zip(a, b):
        assert np.array_equal(_bits(fa), _bits(fb))


def test_keyed_sky_light_equals_indexed(x3d):
    """x3d_scene_load_sky_light by id gives the film of x3d_scene_add_sky_light; the sky is placed
    from the path space's bounds at commit."""
    desc = scenes.sky_scene()
    indexed = x3d.Scene(desc)
    keyed = x3d.Scene()
    for k, m in enumerate(desc.materials):
        keyed.load_material(10 + k, m)
    for k, m in enumerate(desc.meshes):
        keyed.load_geometry(100 + k, m.verts, m.normals, m.tris, material=10 + m.material)
    for k, (mi, rad) in enumerate(desc.lights):
        if mi < 0:
            keyed.load_sky_light(900 + k, rad)
        else:
            keyed.load_light(900 + k, 100 + mi, rad)
    keyed.commit()
    for pt in (x3d.PT_UNIDIRECT_MIS, x3d.PT_BIDIRECT_MIS, x3d.PT_DIRECT):
        a, _ = x3d.Renderer(pt, seed=3).render(keyed, desc.camera, 40, 32, 3)
        b, _ = x3d.Renderer(pt, seed=3).render(indexed, desc.camera, 40, 32, 3)
        assert np.array_equal(_bits(a), _bits(b)) and a[:, :3].max() > 0


@pytest.mark.parametrize("tracer", ["UNIDIRECT_MIS", "DIRECT", "UNIDIRECT_LT1", "UNIDIRECT", "BIDIRECT_LT2",
                                    "BIDIRECT_MIS"])
def test_zero_power_lights_give_black_finite_films(x3d, oracle_mod, tracer):
    import dataclasses

    desc = scenes.multi_light_scene()
    dark = dataclasses.replace(desc, lights=[(m, (0.0, 0.0, 0.0)) for m, _ in desc.lights])
    pt = getattr(x3d, "PT_" + tracer)
    img, _ = x3d.Renderer(pt, seed=2).render(x3d.Scene(dark), dark.camera, 24, 20, 4)
    assert np.isfinite(img).all() and not img[:, :3].any()
    ref = oracle_mod.Oracle(dark).render(pt, 24, 20, 4, seed=2)
    assert np.array_equal(_bits(img), _bits(ref))


def test_negative_radiance_rejected(x3d):
    s = x3d.Scene()
    m = scenes.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), -1)
    mi = s.add_mesh(m.verts, m.normals, m.tris)
    with pytest.raises(x3d.X3DError):
        s.add_area_light(mi, (1.0, -1.0, 1.0))
    with pytest.raises(x3d.X3DError):
        s.add_sky_light((0.5, float("nan"), 0.5))
