"""Writes the golden fixtures under tests/golden/ from the reference's own code.

Run in a container that has the reference checkout and its build (`make -C oracle ref`):

    python tests/golden/make_golden.py

The fixtures are data only: inputs (seeded rays, scene arrays, BSDF directions, and the Cornell
box geometry read from the reference's testdata/cornellbox/*.obj) and the reference's outputs
(flattened BVH dumps, hit mesh / triangle / distance, any-hit answers, BSDF values and samples,
camera rays, layout hashes of the large scenes). They let tests pin the oracle where the
reference build is not available.
"""

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import bindings  # noqa: E402
from x3dcore2_amd import scenes  # noqa: E402

CORNELL_DIR = "/root/reference/testdata/cornellbox"
CORNELL_PARTS = ["left_wall", "right_wall", "back_wall", "ceiling", "floor", "left_sphere", "right_sphere", "light"]


def load_obj(path):
    """Positions, per-vertex normals and triangles of a Wavefront OBJ (v / vn / f a/b/c with the
    normal index of each corner), with one output vertex per distinct (position, normal) pair."""
    pos, nrm, verts, norms, tris, index = [], [], [], [], [], {}
    for line in open(path):
        p = line.split()
        if not p:
            continue
        if p[0] == "v":
            pos.append([float(x) for x in p[1:4]])
        elif p[0] == "vn":
            nrm.append([float(x) for x in p[1:4]])
        elif p[0] == "f":
            corners = []
            for c in p[1:]:
                parts = c.split("/")
                key = (int(parts[0]) - 1, int(parts[2]) - 1)
                if key not in index:
                    index[key] = len(verts)
                    verts.append(pos[key[0]])
                    norms.append(nrm[key[1]])
                corners.append(index[key])
            for k in range(1, len(corners) - 1):
                tris.append([corners[0], corners[k], corners[k + 1]])
    return np.array(verts, np.float32), np.array(norms, np.float32), np.array(tris, np.uint32)


def cornell_scene():
    """The reference's Cornell box (corelib/test/tst_renderer/tst_renderer.cpp:37-115)."""
    mats = [scenes.Material(scenes.OREN_NAYAR, (0.725, 0.710, 0.680), sigma=0.078),
            scenes.Material(scenes.OREN_NAYAR, (0.630, 0.065, 0.050), sigma=0.078),
            scenes.Material(scenes.OREN_NAYAR, (0.140, 0.450, 0.091), sigma=0.078),
            scenes.Material(scenes.COOK_TORR, (0.95, 0.95, 0.95), roughness=0.2, ior=(2.93, 3.0)),
            scenes.Material(scenes.OREN_NAYAR, (0.0, 0.0, 0.0), sigma=0.078)]
    mat_of = {"left_wall": 1, "right_wall": 2, "back_wall": 0, "ceiling": 0, "floor": 0, "left_sphere": 3,
              "right_sphere": 0, "light": 4}
    meshes = []
    for part in CORNELL_PARTS:
        v, n, t = load_obj(os.path.join(CORNELL_DIR, part + ".obj"))
        meshes.append(scenes.Mesh(v, n, t, mat_of[part]))
    rad = tuple(float(np.float32(c) * np.float32(15.0)) for c in (0.911, 0.660, 0.345))
    m = np.eye(4, dtype=np.float32)
    m[0, 3], m[1, 3], m[2, 3] = 0.0, 0.795, 3.4
    cam = scenes.Camera(0.032, 0.035, 4.0 / 3.0, m.T.reshape(-1).copy())
    return scenes.SceneDesc("cornell", mats, meshes, [(len(meshes) - 1, rad)], cam, 800, 600, 1, 5)


def scene_fixture(desc, nrays, seed):
    ref = bindings.Reference(desc)
    b, m, p = ref.bvh()
    rays = scenes.seeded_rays(desc, nrays, seed)
    h = ref.intersect(rays)
    t = np.where(h["geo"] >= 0, h["t"], 2.0).astype(np.float32)
    rng = np.random.default_rng(seed + 1)
    tmin = np.full(nrays, 1e-4, np.float32)
    tmax = (t * rng.choice([0.5, 0.999, 1.001, 2.0], size=nrays)).astype(np.float32)
    occ = ref.occluded(rays, tmin, tmax)
    return dict(bounds=b, meta=m, prims=p, rays=rays, hit_geo=h["geo"], hit_tri=h["tri"], hit_t=h["t"],
                hit_vertex=h["vertex"], hit_normal=h["normal"], occ_tmin=tmin, occ_tmax=tmax, occ=occ)


def layout_sha(ref):
    b, m, p = ref.bvh()
    hsh = hashlib.sha256()
    for a in (b, m, p):
        hsh.update(np.ascontiguousarray(a).tobytes())
    return hsh.hexdigest(), int(b.shape[0]), int(p.shape[0])


def bsdf_fixture():
    desc = scenes.box_scene()
    desc.materials = [scenes.Material(scenes.OREN_NAYAR, (0.73, 0.73, 0.73)),
                      scenes.Material(scenes.OREN_NAYAR, (0.725, 0.710, 0.680), sigma=0.078)]
    rough = (0.05, 0.2, 0.4, 0.7, 0.8)
    desc.materials += [scenes.Material(scenes.COOK_TORR, (0.787, 0.787, 0.787), roughness=r, ior=(2.93, 3.0))
                       for r in rough]
    ref = bindings.Reference(desc)
    rng = np.random.default_rng(2024)
    n = 2000
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    o = rng.normal(size=(n, 3))
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    o *= np.sign((o * nrm).sum(1))[:, None]
    i = rng.normal(size=(n, 3))
    i /= np.linalg.norm(i, axis=1, keepdims=True)
    i *= np.where(rng.random(n) < 0.9, np.sign((i * nrm).sum(1)), 1.0)[:, None]
    nrm, o, i = (x.astype(np.float32) for x in (nrm, o, i))
    out = dict(nrm=nrm, o=o, i=i, roughness=np.array(rough, np.float32))
    for k in range(len(desc.materials)):
        out[f"f{k}"] = ref.bsdf_eval(k, nrm, o, i)
    # tst_material.cpp:40-68 setting: n = (0,0,1), o = normalize(1,1,1), rng seed 13, 1000 samples.
    ns = 1000
    n0 = np.tile(np.array([[0, 0, 1]], np.float32), (ns, 1))
    o0 = np.tile((np.array([[1, 1, 1]]) / np.sqrt(3)).astype(np.float32), (ns, 1))
    out["draws13"] = ref.rng_draws(13, 2 * ns)
    for k in (1, 3):  # oren_nayar 0.078, cook_torr 0.2
        si, sd = ref.bsdf_sample(k, 13, n0, o0)
        out[f"sample_i{k}"] = si
        out[f"sample_d{k}"] = sd
    return out


def camera_fixture():
    """testcamera.cpp:10-51: pinhole (0.032, 0.035, 4/3) rotated 45 degrees about y (mat44_rotate)
    and translated to (1, 2, 3), 1025 x 769 film; a seeded subset of pixels is kept."""
    c, s = np.float32(np.cos(np.pi / 4)), np.float32(np.sin(np.pi / 4))
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float32)
    m[:3, 3] = (1, 2, 3)
    cam = scenes.Camera(0.032, 0.035, 4.0 / 3.0, m.T.reshape(-1).copy())
    ref = bindings.Reference()
    rays = ref.camera_rays(cam, 1025, 769)
    pix = np.random.default_rng(7).choice(1025 * 769, size=4096, replace=False)
    pix[0] = 512 + 384 * 1025
    return dict(transform=cam.transform, pixels=pix.astype(np.int64), rays=rays[pix])


def main():
    if not bindings.reference_available():
        raise SystemExit("oracle/_ref/libx3dref.so is not built (make -C oracle ref)")
    fixtures = {
        "box": (scenes.box_scene(), 4096, 101),
        "soup": (scenes.random_soup(2000, 5), 4096, 102),
        "furnace": (scenes.furnace_sphere(), 2048, 103),
        "cornell": (cornell_scene(), 4096, 104),
    }
    for name, (desc, n, seed) in fixtures.items():
        np.savez_compressed(os.path.join(HERE, f"scene_{name}.npz"), **scene_fixture(desc, n, seed))
        print("wrote", name, desc.num_tris, "tris")
    d = cornell_scene()
    np.savez_compressed(os.path.join(HERE, "cornell_meshes.npz"),
                        **{f"{k}_{part}": getattr(mesh, k) for part, mesh in zip(CORNELL_PARTS, d.meshes)
                           for k in ("verts", "normals", "tris")})
    np.savez_compressed(os.path.join(HERE, "bsdf.npz"), **bsdf_fixture())
    np.savez_compressed(os.path.join(HERE, "camera.npz"), **camera_fixture())
    layouts = {}
    for name, desc in (("spheres", scenes.spheres_scene()), ("indirect", scenes.indirect_scene())):
        sha, nn, npr = layout_sha(bindings.Reference(desc))
        layouts[name] = {"sha256": sha, "nodes": nn, "prims": npr}
        print("layout", name, nn, npr, sha)
    json.dump(layouts, open(os.path.join(HERE, "layouts.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
