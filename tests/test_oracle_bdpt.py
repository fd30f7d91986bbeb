"""The oracle's bidirectional tracer (bidirect_mis_path_tracer, pathtracer.cpp:12-137, 247-304,
351-452, 754-803), pinned without retyping its formulas:

- Closed form. In a closed furnace (every surface a Lambertian reflector of albedo rho and an
  emitter of radiance Le) each strategy of a total length n is an unbiased estimate of the
  radiance carried by paths of n surface vertices, Le * rho^(n-1), and walks never end early, so
  with subpaths of at most M vertices the estimate has expectation
  Le * (1 - rho^(2M+1)) / (1 - rho) per channel. A dropped strategy, a weight that counts the
  cam_plen == 0 strategies, a prefix transport off by one vertex or a projected (cosine-weighted)
  emission would each move this by percent. The fine uv-sphere keeps the tessellation error
  (shading normals vs flat triangles) near 2e-4 relative; its connections have a bounded
  geometry term, so the estimator has no fireflies.
- The reference renderer itself (oracle/_ref, its own bidirect_mis_path_tracer and minstd
  streams) on the box room and on the indirect room with GGX spheres: the images agree in
  8 x 8-pixel blocks within the combined Monte Carlo error.
"""

import numpy as np
import pytest

from x3dcore2_amd import scenes


@pytest.mark.parametrize("max_len", [1, 2, 3, 8])
def test_bdpt_furnace_closed_form(oracle_mod, max_len):
    desc = scenes.furnace_sphere(res=60)
    o = oracle_mod.Oracle(desc)
    img = o.render(oracle_mod.TR_BIDIRECT_MIS, 8, 8, 512, seed=21, prec=1, max_path_len=max_len, threads=8)
    s = img[:, :3].sum(1)
    expected = 3.0 * (1.0 - 0.7 ** (2 * max_len + 1)) / (1.0 - 0.7)
    se = s.std(ddof=1) / np.sqrt(s.size)
    assert abs(s.mean() - expected) < 4 * se + 5e-4 * expected, (s.mean(), expected, se)


def test_bdpt_furnace_fp32_matches_fp64(oracle_mod):
    """The FP32 estimator (the one the GPU reproduces bit for bit) tracks the FP64 one per pixel
    on the same PCG streams; a few pixels differ more where a sample takes another discrete
    branch (a draw against a threshold, a grazing hit) in the two precisions."""
    desc = scenes.furnace_sphere(res=30)
    o = oracle_mod.Oracle(desc)
    a = o.render(oracle_mod.TR_BIDIRECT_MIS, 8, 8, 64, seed=5, prec=0, max_path_len=4, threads=8)
    b = o.render(oracle_mod.TR_BIDIRECT_MIS, 8, 8, 64, seed=5, prec=1, max_path_len=4, threads=8)
    rel = np.abs(a[:, :3] - b[:, :3]) / b[:, :3]
    assert np.median(rel) < 1e-5 and rel.max() < 0.01
    assert abs(a[:, :3].mean() / b[:, :3].mean() - 1) < 1e-3


def _blocks(img, w, h, b):
    x = img.reshape(h // b, b, w // b, b, -1).mean((1, 3))
    return x.reshape(-1, x.shape[-1])


@pytest.mark.parametrize("name", ["box", "indirect"])
def test_bdpt_matches_reference_renderer(oracle_mod, ref_mod, name):
    desc = scenes.box_scene() if name == "box" else scenes.indirect_scene(level=2)
    w = h = 24
    reps, spp = 16, 64
    o = oracle_mod.Oracle(desc)
    imgs = np.stack([o.render(oracle_mod.TR_BIDIRECT_MIS, w, h, spp, seed=300 + s, max_path_len=8, threads=8)
                     [:, :3].astype(np.float64) for s in range(reps)])
    ref = ref_mod.Reference(desc)
    img, taken = ref.render(ref_mod.TR_BIDIRECT_MIS, w, h, 1024, threads=64)
    assert taken == 1024
    ob = np.stack([_blocks(x, w, h, 8) for x in imgs])
    rb = _blocks(img[:, :3].astype(np.float64), w, h, 8)
    m = ob.mean(0)
    se = ob.std(0, ddof=1) / np.sqrt(reps)
    # same estimator => same per-sample variance: scale the oracle's spread to the reference's count
    ref_se = ob.std(0, ddof=1) * np.sqrt(spp / taken)
    z = ((m - rb) / np.sqrt(se ** 2 + ref_se ** 2)).ravel()
    assert np.abs(z).max() < 4.0, z
    assert abs(z.mean()) < 4.0 / np.sqrt(z.size), z.mean()
    rel = np.abs(m.sum(0) - rb.sum(0)) / rb.sum(0)
    assert (rel < 0.01).all(), rel


def test_bdpt_edge_cases(oracle_mod):
    """No emitter: black. A single-vertex camera subpath (max_path_len 1) still sees emission,
    light samples and the one connection strategy."""
    dark = scenes.box_scene()
    dark.lights = []
    img = oracle_mod.Oracle(dark).render(oracle_mod.TR_BIDIRECT_MIS, 8, 8, 4, seed=1)
    assert not img[:, :3].any()
    desc = scenes.box_scene()
    o = oracle_mod.Oracle(desc)
    one = o.render(oracle_mod.TR_BIDIRECT_MIS, 16, 16, 16, seed=2, max_path_len=1)
    full = o.render(oracle_mod.TR_BIDIRECT_MIS, 16, 16, 16, seed=2, max_path_len=8)
    assert one[:, :3].sum() > 0 and full[:, :3].sum() > one[:, :3].sum()
