"""Objects by id on the host (no GPU): trimesh::transform and the keyed-scene rules.

- x3d_transform_mesh equals the reference's own trimesh::transform (oracle/_ref, geometry.cpp:76-87
  with e8util::mat's adjugate inverse, tensor.h:426-535) bit for bit, positions and normals, for
  affine, projective and identity matrices; a singular matrix is refused by both.
- A scene takes objects either by index (x3d_scene_add_*) or by id (x3d_scene_load_*), not both;
  unknown protocols, mixtures by id and bad radiance are refused; nothing is committed yet.
"""

import numpy as np
import pytest

from x3dcore2_amd import scenes


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _col_major(m):
    return np.ascontiguousarray(np.asarray(m, np.float32).T).reshape(-1)


def _matrices():
    rng = np.random.default_rng(4)
    out = [np.eye(4, dtype=np.float32)]
    for _ in range(6):  # affine: rotation-scale-shear + translation
        m = np.eye(4, dtype=np.float32)
        m[:3, :3] = rng.normal(size=(3, 3))
        m[:3, 3] = rng.normal(size=3) * 5
        out.append(m)
    for _ in range(3):  # projective rows (w != 1)
        m = rng.normal(size=(4, 4)).astype(np.float32)
        m[3] = [0.01, -0.02, 0.03, 1.5]
        out.append(m)
    return out


@pytest.mark.parametrize("k", range(10))
def test_transform_matches_reference(x3d, oracle_mod, k):
    if not oracle_mod.reference_available():
        pytest.skip("reference build absent")
    rng = np.random.default_rng(10 + k)
    v = rng.normal(size=(500, 3)).astype(np.float32) * 3
    n = rng.normal(size=(500, 3)).astype(np.float32)
    v[:5] = [[0, 0, 0], [-0.0, 1, 2], [1e-30, -1e-30, 5], [3, -0.0, -0.0], [1e3, 2e3, -4e3]]
    t = _col_major(_matrices()[k])
    a = x3d.transform_mesh(t, v, n)
    b = oracle_mod.ref_transform_mesh(t, v, n)
    assert b is not None
    assert np.array_equal(_bits(a[0]), _bits(b[0])) and np.array_equal(_bits(a[1]), _bits(b[1]))


def test_singular_transform_refused(x3d, oracle_mod):
    m = np.eye(4, dtype=np.float32)
    m[2, 2] = 0.0
    v = np.ones((3, 3), np.float32)
    with pytest.raises(x3d.X3DError):
        x3d.transform_mesh(_col_major(m), v, v)
    if oracle_mod.reference_available():
        assert oracle_mod.ref_transform_mesh(_col_major(m), v, v) is None


def test_keyed_rules(x3d):
    desc = scenes.box_scene()
    mesh = desc.meshes[0]
    s = x3d.Scene()
    s.load_material(7, desc.materials[0])
    s.load_geometry(3, mesh.verts, mesh.normals, mesh.tris, material=7)
    with pytest.raises(x3d.X3DError):            # by index after by id
        s.add_mesh(mesh.verts, mesh.normals, mesh.tris)
    with pytest.raises(x3d.X3DError):
        s.unload(2, 3)                            # camera protocol: not a scene object
    with pytest.raises(x3d.X3DError):
        s.load_material(8, scenes.Material(scenes.MIXTURE, sub0=0, sub1=0, ratio=0.5))
    with pytest.raises(x3d.X3DError):
        s.load_light(9, 3, (1.0, -1.0, 1.0))
    with pytest.raises(x3d.X3DError):             # not committed
        s.geometry_obj_id(0)
    s.unload(x3d.OBJ_GEOMETRY, 12345)             # unknown ids are ignored
    t = x3d.Scene()
    t.add_material(desc.materials[0])
    with pytest.raises(x3d.X3DError):             # by id after by index
        t.load_material(1, desc.materials[0])
    bad = np.eye(4, dtype=np.float32)
    bad[0, 0] = np.nan
    with pytest.raises(x3d.X3DError):
        s.load_geometry(4, mesh.verts, mesh.normals, mesh.tris, transform=_col_major(bad))
