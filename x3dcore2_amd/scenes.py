"""Seeded procedural scenes for the renderer's workloads (flat world-space triangle meshes).

A scene is plain data: materials, meshes (positions, shading normals, triangles, material
index), area lights and a pinhole camera. The same arrays are loaded into the HIP renderer and
into the CPU checkers, so every backend sees bit-identical inputs.

Coordinates: the room is the unit cube x, y, z in [0, 1] with the open side at z = 1; y is up.
The camera sits in front of the opening and looks down -z, as the reference's pinhole camera
does in its own frame (corelib/src/camera.cpp:38-48).
"""

from __future__ import annotations

import dataclasses
import functools
import math
from typing import List, Optional, Tuple

import numpy as np

# Material kinds, matching include/x3dpt.h.
OREN_NAYAR = 0  # reference oren_nayar (material.cpp:108-169); sigma = 0 is Lambertian
COOK_TORR = 1  # reference cook_torr (material.cpp:171-252), GGX
MIXTURE = 2  # reference mat_mixture (material.cpp:78-106)

GLOSSY_IOR = (2.93, 3.0)  # the reference's glossy metal (corelib/test/tst_renderer.cpp:49-50)


@dataclasses.dataclass
class Material:
    kind: int = OREN_NAYAR
    albedo: Tuple[float, float, float] = (0.8, 0.8, 0.8)
    sigma: float = 0.0
    roughness: float = 0.0
    ior: Tuple[float, float] = (1.0, 0.0)
    sub0: int = -1
    sub1: int = -1
    ratio: float = 0.5


@dataclasses.dataclass
class Mesh:
    verts: np.ndarray  # float32 [n, 3]
    normals: np.ndarray  # float32 [n, 3]
    tris: np.ndarray  # uint32 [m, 3]
    material: int = -1


@dataclasses.dataclass
class Camera:
    sensor: float
    focal: float
    aspect: float
    transform: np.ndarray  # float32 [16], column-major 4x4 (element (i, j) at 4*j + i)

    def packed(self) -> np.ndarray:
        return np.concatenate([[self.sensor, self.focal, self.aspect], self.transform]).astype(np.float32)


@dataclasses.dataclass
class SceneDesc:
    name: str
    materials: List[Material]
    meshes: List[Mesh]
    lights: List[Tuple[int, Tuple[float, float, float]]]
    camera: Camera
    width: int = 256
    height: int = 256
    spp: int = 16
    max_depth: int = 5

    @property
    def num_tris(self) -> int:
        return int(sum(m.tris.shape[0] for m in self.meshes))


def translate_camera(x: float, y: float, z: float, sensor=0.036, focal=0.035, aspect=1.0) -> Camera:
    m = np.eye(4, dtype=np.float32)
    m[0, 3], m[1, 3], m[2, 3] = x, y, z
    return Camera(sensor, focal, aspect, m.T.reshape(-1).copy())


# ------------------------------------------------------------------------------------------
# Primitives
# ------------------------------------------------------------------------------------------
def quad(p0, p1, p2, p3, normal, material) -> Mesh:
    v = np.array([p0, p1, p2, p3], dtype=np.float32)
    n = np.repeat(np.array([normal], dtype=np.float32), 4, axis=0)
    t = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    return Mesh(v, n, t, material)


def box(center, size, angle_y, material) -> Mesh:
    """Axis-aligned box rotated by angle_y (radians) about the vertical axis through center."""
    cx, cy, cz = center
    sx, sy, sz = size
    c, s = math.cos(angle_y), math.sin(angle_y)

    def rot(p):
        x, y, z = p
        return (c * x + s * z, y, -s * x + c * z)

    faces = [  # (normal, 4 corners as unit-cube offsets in [-1, 1])
        ((1, 0, 0), [(1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1)]),
        ((-1, 0, 0), [(-1, -1, 1), (-1, 1, 1), (-1, 1, -1), (-1, -1, -1)]),
        ((0, 1, 0), [(-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1)]),
        ((0, -1, 0), [(-1, -1, 1), (-1, -1, -1), (1, -1, -1), (1, -1, 1)]),
        ((0, 0, 1), [(-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]),
        ((0, 0, -1), [(1, -1, -1), (-1, -1, -1), (-1, 1, -1), (1, 1, -1)]),
    ]
    verts, norms, tris = [], [], []
    for nrm, corners in faces:
        base = len(verts)
        for o in corners:
            p = rot((o[0] * sx / 2, o[1] * sy / 2, o[2] * sz / 2))
            verts.append((p[0] + cx, p[1] + cy, p[2] + cz))
            norms.append(rot(nrm))
        tris += [(base, base + 1, base + 2), (base, base + 2, base + 3)]
    return Mesh(np.array(verts, np.float32), np.array(norms, np.float32), np.array(tris, np.uint32), material)


def icosphere(level: int) -> Tuple[np.ndarray, np.ndarray]:
    """Unit icosphere with 20 * 4**level faces (vertices on the sphere, shared)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, dtype=np.int64)
    for _ in range(level):
        nv = verts.shape[0]
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        e = np.sort(e, axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mids = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        verts = np.concatenate([verts, mids])
        nf = faces.shape[0]
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        ab, bc, ca = inv[:nf] + nv, inv[nf:2 * nf] + nv, inv[2 * nf:] + nv
        faces = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                                np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    return verts, faces


def _value_noise(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise in [-1, 1] on the integer lattice, hashed from `seed`."""
    base = np.floor(p).astype(np.int64)
    f = p - base
    f = f * f * (3 - 2 * f)

    def h(ix, iy, iz):
        x = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        x = (x ^ (x >> 13)) * 1274126177
        x = x ^ (x >> 16)
        return ((x & 0xFFFF).astype(np.float64) / 65535.0) * 2.0 - 1.0

    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1])
                     * (f[:, 2] if dz else 1 - f[:, 2]))
                out += w * h(base[:, 0] + dx, base[:, 1] + dy, base[:, 2] + dz)
    return out


def vertex_normals(verts: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """Area-weighted vertex normals of a closed mesh."""
    fn = np.cross(verts[faces[:, 1]] - verts[faces[:, 0]], verts[faces[:, 2]] - verts[faces[:, 0]])
    n = np.zeros_like(verts)
    for k in range(3):
        np.add.at(n, faces[:, k], fn)
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def displaced_sphere(center, radius, level, amplitude, seed, material) -> Mesh:
    verts, faces = icosphere(level)
    disp = 1.0 + amplitude * _value_noise(verts * 4.0 + 17.0 * (seed % 7), seed)
    v = verts * (radius * disp)[:, None] + np.asarray(center)[None, :]
    n = vertex_normals(v, faces)
    return Mesh(v.astype(np.float32), n.astype(np.float32), faces.astype(np.uint32), material)


# ------------------------------------------------------------------------------------------
# Workloads
# ------------------------------------------------------------------------------------------
def _room(mats: List[Material], meshes: List[Mesh], ceiling_light=True):
    """Five inward quads of the unit room; returns the material indices (white, red, green)."""
    white = len(mats)
    mats.append(Material(OREN_NAYAR, (0.73, 0.73, 0.73)))
    red = len(mats)
    mats.append(Material(OREN_NAYAR, (0.63, 0.07, 0.05)))
    green = len(mats)
    mats.append(Material(OREN_NAYAR, (0.14, 0.45, 0.09)))
    meshes.append(quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), white))  # floor
    meshes.append(quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), (0, -1, 0), white))  # ceiling
    meshes.append(quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), white))  # back
    meshes.append(quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), red))  # left
    meshes.append(quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0), green))  # right
    return white, red, green


def _ceiling_emitter(mats, meshes, lights, radiance=(17.0, 12.0, 4.0), half=0.125, y=0.999):
    black = len(mats)
    mats.append(Material(OREN_NAYAR, (0.0, 0.0, 0.0)))
    lo, hi = 0.5 - half, 0.5 + half
    meshes.append(quad((lo, y, lo), (hi, y, lo), (hi, y, hi), (lo, y, hi), (0, -1, 0), black))
    lights.append((len(meshes) - 1, radiance))


def box_scene(width=256, height=256, spp=16, max_depth=5) -> SceneDesc:
    """Lambertian room with two blocks rotated +-17 degrees and a 0.25 x 0.25 ceiling emitter
    of radiance (17, 12, 4): 36 triangles."""
    mats: List[Material] = []
    meshes: List[Mesh] = []
    lights = []
    white, _, _ = _room(mats, meshes)
    meshes.append(box((0.33, 0.15, 0.62), (0.3, 0.3, 0.3), math.radians(-17.0), white))
    meshes.append(box((0.66, 0.3, 0.35), (0.3, 0.6, 0.3), math.radians(17.0), white))
    _ceiling_emitter(mats, meshes, lights)
    return SceneDesc("box", mats, meshes, lights, translate_camera(0.5, 0.5, 2.0), width, height, spp, max_depth)


def fan(points, normal, material) -> Mesh:
    """Planar polygon as a triangle fan from its first corner (triangles of unequal area unless
    the polygon is regular)."""
    v = np.array(points, dtype=np.float32)
    n = np.repeat(np.array([normal], dtype=np.float32), v.shape[0], axis=0)
    t = np.array([[0, k, k + 1] for k in range(1, v.shape[0] - 1)], dtype=np.uint32)
    return Mesh(v, n, t, material)


def multi_light_scene(width=256, height=256, spp=16, max_depth=5, unequal=True) -> SceneDesc:
    """The box room with three emitters of unequal power and area: the 0.25 x 0.25 ceiling quad
    (17, 12, 4), a 0.3 x 0.15 panel in front of the back wall facing +z (2, 4, 6) and a
    pentagon in front of the left wall facing +x (6, 6, 3). With unequal=True the pentagon is a
    fan of three triangles of different areas; with unequal=False it is split into equal halves
    (two triangles of one quad), the case where the reference's choice of emitter triangle by
    index (geometry.cpp:47-48, 63) is unbiased."""
    mats: List[Material] = []
    meshes: List[Mesh] = []
    lights = []
    white, _, _ = _room(mats, meshes)
    meshes.append(box((0.33, 0.15, 0.62), (0.3, 0.3, 0.3), math.radians(-17.0), white))
    meshes.append(box((0.66, 0.3, 0.35), (0.3, 0.6, 0.3), math.radians(17.0), white))
    _ceiling_emitter(mats, meshes, lights)
    black = len(mats) - 1
    z = 0.002
    meshes.append(quad((0.2, 0.55, z), (0.5, 0.55, z), (0.5, 0.7, z), (0.2, 0.7, z), (0, 0, 1), black))
    lights.append((len(meshes) - 1, (2.0, 4.0, 6.0)))
    x = 0.002
    if unequal:
        pent = [(x, 0.45, 0.3), (x, 0.42, 0.55), (x, 0.6, 0.72), (x, 0.78, 0.6), (x, 0.72, 0.35)]
        meshes.append(fan(pent, (1, 0, 0), black))
    else:
        meshes.append(quad((x, 0.45, 0.3), (x, 0.45, 0.7), (x, 0.75, 0.7), (x, 0.75, 0.3), (1, 0, 0), black))
    lights.append((len(meshes) - 1, (6.0, 6.0, 3.0)))
    return SceneDesc("multi_light" if unequal else "multi_light_eq", mats, meshes, lights,
                     translate_camera(0.5, 0.5, 2.0), width, height, spp, max_depth)


SKY_RADIANCE = (0.529, 0.808, 0.922)  # the reference's default sky (resource.cpp:518-523)


def sky_scene(width=256, height=256, spp=16, max_depth=5, area_light=True) -> SceneDesc:
    """An open scene under the sky light: a ground quad in the plane z = 0 facing +z (the sky's
    zenith), a block, a tall block and a GGX sphere on it, seen from above; optionally a small
    area emitter facing the ground as a second light of another power."""
    mats: List[Material] = [Material(OREN_NAYAR, (0.6, 0.6, 0.55)), Material(OREN_NAYAR, (0.7, 0.3, 0.2)),
                            Material(COOK_TORR, (0.8, 0.8, 0.8), roughness=0.3, ior=GLOSSY_IOR),
                            Material(OREN_NAYAR, (0.0, 0.0, 0.0))]
    meshes: List[Mesh] = [quad((-0.5, -0.5, 0), (1.5, -0.5, 0), (1.5, 1.5, 0), (-0.5, 1.5, 0), (0, 0, 1), 0)]
    m = box((0.3, 0.35, 0.15), (0.3, 0.3, 0.3), math.radians(25.0), 1)
    meshes.append(m)
    meshes.append(box((0.7, 0.6, 0.3), (0.2, 0.2, 0.6), math.radians(-10.0), 0))
    meshes.append(displaced_sphere((0.45, 0.75, 0.15), 0.15, 3, 0.0, 5, 2))
    lights = [(-1, SKY_RADIANCE)]
    if area_light:
        meshes.append(quad((0.6, 0.1, 0.8), (0.6, 0.3, 0.8), (0.8, 0.3, 0.8), (0.8, 0.1, 0.8), (0, 0, -1), 3))
        lights.append((len(meshes) - 1, (8.0, 8.0, 6.0)))
    return SceneDesc("sky", mats, meshes, lights, translate_camera(0.5, 0.5, 2.0), width, height, spp, max_depth)


def sphere_levels(rng: np.random.Generator, count: int, target_tris: int) -> np.ndarray:
    """Subdivision levels in {3, 4, 5} whose face total is closest to target_tris (seeded)."""
    levels = rng.choice([3, 4, 5], size=count, p=[0.48, 0.30, 0.22])
    faces = lambda lv: int(np.sum(20 * 4 ** lv.astype(np.int64)))
    # Greedy seeded adjustment towards the target.
    order = rng.permutation(count)
    for i in order:
        cur = faces(levels)
        if abs(cur - target_tris) < 20 * 4 ** 3:
            break
        if cur > target_tris and levels[i] > 3:
            levels[i] -= 1
        elif cur < target_tris and levels[i] < 5:
            levels[i] += 1
    return levels


def spheres_scene(width=1024, height=1024, spp=64, max_depth=8, count=200, target_tris=1_300_000,
                  seed=0x5EED0002) -> SceneDesc:
    """Room plus `count` displaced icospheres (about 1.3M triangles at the default count): centres
    uniform in the room, radii log-uniform in [0.02, 0.15], displacement 0-5% from value noise,
    per-object GGX roughness U[0.05, 0.8] and albedo U[0.1, 0.9]^3."""
    rng = np.random.default_rng(seed)
    mats: List[Material] = []
    meshes: List[Mesh] = []
    lights = []
    _room(mats, meshes)
    levels = sphere_levels(rng, count, target_tris)
    for k in range(count):
        r = float(math.exp(rng.uniform(math.log(0.02), math.log(0.15))))
        c = rng.uniform(r, 1.0 - r, size=3)
        amp = float(rng.uniform(0.0, 0.05))
        rough = float(rng.uniform(0.05, 0.8))
        alb = tuple(float(a) for a in rng.uniform(0.1, 0.9, size=3))
        mats.append(Material(COOK_TORR, alb, roughness=rough, ior=GLOSSY_IOR))
        meshes.append(displaced_sphere(c, r, int(levels[k]), amp, seed + k, len(mats) - 1))
    _ceiling_emitter(mats, meshes, lights)
    return SceneDesc("spheres", mats, meshes, lights, translate_camera(0.5, 0.5, 2.0), width, height, spp,
                     max_depth)


def indirect_scene(width=512, height=512, spp=32, max_path_len=6, level=5) -> SceneDesc:
    """Room whose emitter faces the ceiling above a 0.5 x 0.5 occluding slab, so light reaches
    the scene only after bouncing; four GGX icospheres of roughness 0.05 / 0.2 / 0.4 / 0.7."""
    mats: List[Material] = []
    meshes: List[Mesh] = []
    lights = []
    white, _, _ = _room(mats, meshes)
    meshes.append(box((0.5, 0.85, 0.5), (0.5, 0.02, 0.5), 0.0, white))  # occluder below the emitter
    black = len(mats)
    mats.append(Material(OREN_NAYAR, (0.0, 0.0, 0.0)))
    lo, hi, y = 0.375, 0.625, 0.87
    meshes.append(quad((lo, y, lo), (lo, y, hi), (hi, y, hi), (hi, y, lo), (0, 1, 0), black))
    lights.append((len(meshes) - 1, (17.0, 12.0, 4.0)))
    for k, (rough, cx) in enumerate(zip((0.05, 0.2, 0.4, 0.7), (0.17, 0.39, 0.61, 0.83))):
        mats.append(Material(COOK_TORR, (0.9, 0.9, 0.9), roughness=rough, ior=GLOSSY_IOR))
        meshes.append(displaced_sphere((cx, 0.12, 0.5), 0.1, level, 0.0, 7 + k, len(mats) - 1))
    desc = SceneDesc("indirect", mats, meshes, lights, translate_camera(0.5, 0.5, 2.0), width, height, spp,
                     max_path_len)
    return desc


ADJOINT_GROUPS = 16


def adjoint_truth(seed=0x5EED0004) -> Tuple[np.ndarray, np.ndarray]:
    """Seeded ground-truth parameters of the inverse-rendering workload: albedo [16, 3] in
    [0.15, 0.85] and GGX roughness [16] in [0.2, 0.7]."""
    rng = np.random.default_rng(seed)
    return (rng.uniform(0.15, 0.85, size=(ADJOINT_GROUPS, 3)).astype(np.float32),
            rng.uniform(0.2, 0.7, size=ADJOINT_GROUPS).astype(np.float32))


def adjoint_scene(albedo=None, roughness=None, width=256, height=256, spp=32, max_depth=4) -> SceneDesc:
    """The box room cut into 16 parametric surfaces: floor (front and back halves), ceiling,
    back, left and right walls, and the five visible faces of each block (the blocks' bottoms
    share the front floor's parameters). Surface m is a 50/50 mixture of a Lambertian and a GGX
    lobe (reference mat_mixture of oren_nayar and cook_torr) that share albedo[m]; the GGX lobe
    has roughness[m]. Material ids: 3m Lambertian, 3m+1 GGX, 3m+2 the mixture; the emitter's
    black Lambertian is 48."""
    albedo = np.full((ADJOINT_GROUPS, 3), 0.5, np.float32) if albedo is None else np.asarray(albedo, np.float32)
    roughness = np.full(ADJOINT_GROUPS, 0.4, np.float32) if roughness is None else np.asarray(roughness, np.float32)
    mats: List[Material] = []
    for m in range(ADJOINT_GROUPS):
        a = tuple(float(x) for x in albedo[m])
        mats.append(Material(OREN_NAYAR, a))
        mats.append(Material(COOK_TORR, a, roughness=float(roughness[m]), ior=GLOSSY_IOR))
        mats.append(Material(MIXTURE, sub0=3 * m, sub1=3 * m + 1, ratio=0.5))
    mix = lambda m: 3 * m + 2
    meshes = [
        quad((0, 0, 0.5), (1, 0, 0.5), (1, 0, 1), (0, 0, 1), (0, 1, 0), mix(0)),  # floor, front half
        quad((0, 0, 0), (1, 0, 0), (1, 0, 0.5), (0, 0, 0.5), (0, 1, 0), mix(1)),  # floor, back half
        quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), (0, -1, 0), mix(2)),  # ceiling
        quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), mix(3)),  # back
        quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), mix(4)),  # left
        quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0), mix(5)),  # right
    ]
    # box() face order: +x, -x, +y, -y, +z, -z (4 vertices, 2 triangles each)
    for base, (center, size, ang) in ((6, ((0.33, 0.15, 0.62), (0.3, 0.3, 0.3), -17.0)),
                                      (11, ((0.66, 0.3, 0.35), (0.3, 0.6, 0.3), 17.0))):
        b = box(center, size, math.radians(ang), 0)
        groups = [base, base + 1, base + 2, 0, base + 3, base + 4]
        for f, g in enumerate(groups):
            meshes.append(Mesh(b.verts[4 * f:4 * f + 4].copy(), b.normals[4 * f:4 * f + 4].copy(),
                               np.array([[0, 1, 2], [0, 2, 3]], np.uint32), mix(g)))
    lights = []
    _ceiling_emitter(mats, meshes, lights)
    return SceneDesc("adjoint", mats, meshes, lights, translate_camera(0.5, 0.5, 2.0), width, height, spp,
                     max_depth)


def tie_gradient(grad: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Per-material gradient [num_materials, 4] of an adjoint_scene -> gradient with respect to
    the 16 shared albedos [16, 3] and roughnesses [16]."""
    g = np.asarray(grad, np.float64)
    ga = np.stack([g[3 * m, :3] + g[3 * m + 1, :3] for m in range(ADJOINT_GROUPS)])
    gr = np.array([g[3 * m + 1, 3] for m in range(ADJOINT_GROUPS)])
    return ga, gr


def set_adjoint_params(scene, albedo, roughness, update_sampling=False):
    """Writes the 16 shared parameters into an adjoint_scene's material table (a device Scene
    or an oracle, both with set_material_params(first, albedo, roughness, update_sampling))."""
    a = np.repeat(np.asarray(albedo, np.float32), 3, axis=0).reshape(-1, 3)  # ON, GGX, mixture
    r = np.repeat(np.asarray(roughness, np.float32), 3)
    scene.set_material_params(0, a, r, update_sampling)


@functools.lru_cache(maxsize=None)
def _unit_icosphere(level: int) -> Tuple[np.ndarray, np.ndarray]:
    v, f = icosphere(level)
    return v.astype(np.float32), f.astype(np.uint32)


def batch_scenes(count=64, seed=0x5EED0003, min_tris=1_000, max_tris=2_000_000, max_depth=6,
                 resolutions=(128, 256, 512, 1024), spps=(4, 16, 64)) -> List[SceneDesc]:
    """Seeded batch of scenes: the room and ceiling emitter plus instanced icospheres (unit
    icospheres of subdivision level 1-7, scaled and placed, flattened to world space) and boxes,
    total triangle count log-uniform in [min_tris, max_tris]; each object Lambertian or GGX
    (roughness U[0.05, 0.8]); square resolution and spp drawn from the given sets."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        target = float(math.exp(rng.uniform(math.log(min_tris), math.log(max_tris))))
        mats: List[Material] = []
        meshes: List[Mesh] = []
        lights = []
        _room(mats, meshes)
        _ceiling_emitter(mats, meshes, lights)
        tris = 12
        while tris + 12 <= target:
            left = target - tris
            alb = tuple(float(a) for a in rng.uniform(0.1, 0.9, size=3))
            if rng.random() < 0.5:
                mats.append(Material(OREN_NAYAR, alb))
            else:
                mats.append(Material(COOK_TORR, alb, roughness=float(rng.uniform(0.05, 0.8)), ior=GLOSSY_IOR))
            mat = len(mats) - 1
            r = float(math.exp(rng.uniform(math.log(0.02), math.log(0.15))))
            c = rng.uniform(r, 1.0 - r, size=3)
            top = int(min(7, math.floor(math.log(max(left, 20) / 20.0, 4))))
            if top < 1 or rng.random() < 0.25:
                meshes.append(box(tuple(c), (2 * r, 2 * r, 2 * r), float(rng.uniform(0, math.pi)), mat))
                tris += 12
                continue
            level = int(rng.integers(max(1, top - 1), top + 1))
            v, f = _unit_icosphere(level)
            meshes.append(Mesh((v * np.float32(r) + c.astype(np.float32)).astype(np.float32), v.copy(), f, mat))
            tris += f.shape[0]
        res = int(rng.choice(resolutions))
        spp = int(rng.choice(spps))
        out.append(SceneDesc(f"batch{k}", mats, meshes, lights, translate_camera(0.5, 0.5, 2.0), res, res, spp,
                             max_depth))
    return out


def furnace_sphere(radius=10.0, res=30, albedo=0.7, radiance=1.0) -> SceneDesc:
    """Inward uv-sphere that is both an Oren-Nayar(sigma=0) reflector and an area light, the
    closed-form furnace of corelib/test/tst_pathtracer/tst_pathtracer.cpp:24-56
    (expected radiance L / (1 - albedo) per channel). Geometry follows geometry.cpp:162-241."""
    verts, norms = [], []
    verts.append((0.0, 0.0, -radius))
    norms.append((0.0, 0.0, 1.0))
    for j in range(1, res - 1):
        u = np.float32(j) / np.float32(res)
        phi = (np.float32(1) - u) * np.float32(math.pi)
        for i in range(res):
            v = np.float32(i) / np.float32(res)
            theta = v * np.float32(2 * math.pi)
            x = np.float32(math.sin(phi) * math.cos(theta))
            y = np.float32(math.sin(phi) * math.sin(theta))
            z = np.float32(math.cos(phi))
            verts.append((radius * x, radius * y, radius * z))
            norms.append((-x, -y, -z))
    verts.append((0.0, 0.0, radius))
    norms.append((0.0, 0.0, -1.0))
    tris = []
    for i in range(1, res + 1):
        tris.append((i % res + 1, i, 0))
    for j in range(1, res - 2):
        for i in range(1, res + 1):
            tris.append((i % res + 1 + j * res, i + j * res, i + (j - 1) * res))
            tris.append((i % res + 1 + (j - 1) * res, i % res + 1 + j * res, i + (j - 1) * res))
    for i in range(1, res + 1):
        tris.append((i % res + 1 + (res - 3) * res, 1 + (res - 2) * res, i + (res - 3) * res))
    mats = [Material(OREN_NAYAR, (albedo, albedo, albedo))]
    mesh = Mesh(np.array(verts, np.float32), np.array(norms, np.float32), np.array(tris, np.uint32), 0)
    return SceneDesc("furnace", mats, [mesh], [(0, (radiance, radiance, radiance))],
                     translate_camera(0.0, 0.0, 0.0, aspect=1.0), 16, 16, 64, 5)


def furnace_box(half=2.5, albedo=0.7, radiance=1.0) -> SceneDesc:
    """Closed cube (12 flat triangles, inward normals) that is both a Lambertian reflector and an
    area light, with the camera inside. Shading normals equal the geometric ones, so walks never
    end on a back-facing hit and the radiance carried by paths of at most n vertices is exactly
    radiance * (1 - albedo^n) / (1 - albedo) per channel."""
    m = box((0.5, 0.5, 0.5), (2 * half, 2 * half, 2 * half), 0.0, 0)
    m.normals = -m.normals
    mats = [Material(OREN_NAYAR, (albedo, albedo, albedo))]
    return SceneDesc("furnace_box", mats, [m], [(0, (radiance, radiance, radiance))],
                     translate_camera(0.5, 0.5, 2.0), 16, 16, 64, 5)


def random_soup(num_tris: int, seed: int, extent=1.0) -> SceneDesc:
    """Random triangles in a cube (for hit-id parity on non-axis-aligned geometry)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, extent, size=(num_tris, 1, 3))
    v = (c + rng.normal(0, 0.05 * extent, size=(num_tris, 3, 3))).reshape(-1, 3).astype(np.float32)
    fn = np.cross(v[1::3] - v[0::3], v[2::3] - v[0::3])
    fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), 1e-12)
    n = np.repeat(fn, 3, axis=0).astype(np.float32)
    t = np.arange(3 * num_tris, dtype=np.uint32).reshape(-1, 3)
    mats = [Material(OREN_NAYAR, (0.5, 0.5, 0.5))]
    return SceneDesc("soup", mats, [Mesh(v, n, t, 0)], [], translate_camera(0.5, 0.5, 2.0))


def tie_scene(seed: int = 5) -> SceneDesc:
    """Geometry where hits tie: a 10 x 10 grid of quads on the plane z = 0.5 repeated in 8 meshes
    (identical triangles in different leaves, hit at equal distances), 400 copies of one
    triangle in one mesh (leaves whose boxes and centroids coincide) and 1000 random triangles."""
    rng = np.random.default_rng(seed)
    meshes = []
    g = np.linspace(0.1, 0.9, 11, dtype=np.float32)
    quads = []
    for i in range(10):
        for j in range(10):
            a, b = (g[i], g[j]), (g[i + 1], g[j + 1])
            quads += [(a[0], a[1]), (b[0], a[1]), (b[0], b[1]), (a[0], a[1]), (b[0], b[1]), (a[0], b[1])]
    gv = np.array([(x, y, 0.5) for x, y in quads], np.float32)
    gn = np.tile(np.array([0, 0, 1], np.float32), (gv.shape[0], 1))
    gt = np.arange(gv.shape[0], dtype=np.uint32).reshape(-1, 3)
    for _ in range(8):
        meshes.append(Mesh(gv.copy(), gn.copy(), gt.copy(), 0))
    one = np.array([[0.3, 0.2, 0.7], [0.8, 0.3, 0.7], [0.5, 0.8, 0.75]], np.float32)
    dv = np.tile(one, (400, 1))
    fn = np.cross(one[1] - one[0], one[2] - one[0])
    fn = (fn / np.linalg.norm(fn)).astype(np.float32)
    meshes.append(Mesh(dv, np.tile(fn, (dv.shape[0], 1)), np.arange(dv.shape[0], dtype=np.uint32).reshape(-1, 3), 0))
    soup = random_soup(1000, seed)
    meshes.append(soup.meshes[0])
    mats = [Material(OREN_NAYAR, (0.5, 0.5, 0.5))]
    return SceneDesc("ties", mats, meshes, [], translate_camera(0.5, 0.5, 2.0))


def seeded_rays(desc: SceneDesc, n: int, seed: int, inside=True) -> np.ndarray:
    """Rays as float32 [n, 6] (origin, unit direction): origins uniform in the room (or
    around the scene bounds), directions uniform on the sphere, plus a share of camera-like rays
    and rays aimed at triangle vertices and edges (where ties between triangles occur)."""
    rng = np.random.default_rng(seed)
    allv = np.concatenate([m.verts for m in desc.meshes])
    lo, hi = allv.min(0), allv.max(0)
    pad = 0.05 * (hi - lo + 1e-6)
    o = rng.uniform(lo + pad, hi - pad, size=(n, 3)) if inside else rng.uniform(lo - 2 * pad, hi + 2 * pad,
                                                                                    size=(n, 3))
    d = rng.normal(size=(n, 3))
    # A quarter of the rays aim at vertices or edge midpoints of random triangles.
    k = n // 4
    mesh_ids = rng.integers(0, len(desc.meshes), size=k)
    tgt = np.empty((k, 3))
    for j, mi in enumerate(mesh_ids):
        m = desc.meshes[mi]
        t = m.tris[rng.integers(0, m.tris.shape[0])]
        w = rng.choice(4)
        if w < 3:
            tgt[j] = m.verts[t[w]]
        else:
            tgt[j] = 0.5 * (m.verts[t[0]].astype(np.float64) + m.verts[t[1]])
    d[:k] = tgt - o[:k]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)
