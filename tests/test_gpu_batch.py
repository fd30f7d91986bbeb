"""Batched rendering through the C ABI: every job's film must be bit-identical to rendering that
job alone with the same renderer (and so to the oracle), on first capture and on graph replay.
Plain form: whatever the workgroup size, stack depth, tile order and split between the shared
launch and jobs rendered on their own in the wavefront form. Wavefront form (all jobs' paths in
shared generations): whatever the generation size (one generation for the whole batch, jobs split
over several generations, generations of a single film's pixels), workgroup size and stack depth."""

import numpy as np
import pytest

from x3dcore2_amd import scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def batch(x3d):
    descs = scenes.batch_scenes(count=9, seed=77, min_tris=200, max_tris=40000, resolutions=(32, 48, 64),
                                spps=(1, 2, 5))
    descs.append(scenes.box_scene())
    descs[-1].width = descs[-1].height = 40
    descs[-1].spp = 3
    print("[gpu] batch scenes:", [(d.num_tris, d.width, d.spp) for d in descs], flush=True)
    built = [x3d.Scene(d) for d in descs]
    items = [(built[k], d.camera, d.width, d.height, d.spp) for k, d in enumerate(descs)]
    items.append((built[0], descs[0].camera, 33, 17, 2))  # a scene shared by two jobs, ragged size
    return descs, built, items


@pytest.mark.parametrize("kw", [dict(), dict(workgroup_size=64, stack_depth=8, sort_key=2),
                                dict(workgroup_size=128, sort_key=1),
                                dict(batch_split=3000), dict(batch_split=1, queue_capacity=5000)])
def test_batch_films_identical_to_single_renders(x3d, batch, kw):
    descs, built, items = batch
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=6, seed=9, **kw)
    singles = [r.render(s, cam, w, h, spp)[0] for (s, cam, w, h, spp) in items]
    for rep in range(2):  # capture, then replay
        print(f"[gpu] batch {kw} pass {rep}", flush=True)
        films, st, lat = r.render_batch(items)
        for k, (f, ref) in enumerate(zip(films, singles)):
            assert np.array_equal(_bits(f), _bits(ref)), (k, np.abs(f - ref).max())
        assert st.num_samples == sum(w * h * spp for (_, _, w, h, spp) in items)
        assert st.closest_queries >= sum(w * h for (_, _, w, h, _) in items)
        split = kw.get("batch_split", 1 << 22)
        assert st.alone_jobs == sum(w * h * spp >= split for (_, _, w, h, spp) in items)
        assert all(0 < x <= st.kernel_ms * 1.01 + 0.05 for x in lat), (lat, st.kernel_ms)


def test_batch_matches_oracle(x3d, oracle_mod, batch):
    descs, built, items = batch
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=6, seed=3)
    films, _, _ = r.render_batch(items[:3] + items[-2:])
    for k, (f, d) in enumerate(zip(films, descs[:3] + descs[-1:])):
        ref = oracle_mod.Oracle(d).render(oracle_mod.TR_UNIDIRECT_MIS, d.width, d.height, d.spp, seed=3, max_depth=6)
        assert np.array_equal(_bits(f), _bits(ref)), k


def test_batch_autotune_and_errors(x3d, batch):
    descs, built, items = batch
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=6, seed=9)
    ref, _, _ = r.render_batch(items)
    rep = r.autotune_batch(items)
    assert rep.candidates == 12 and rep.best_ms > 0 and rep.workgroup_size in (64, 128, 256)
    films, _, _ = r.render_batch(items)
    for a, b in zip(films, ref):
        assert np.array_equal(_bits(a), _bits(b))
    with pytest.raises(x3d.X3DError):
        x3d.Renderer(x3d.PT_BIDIRECT_MIS).render_batch(items[:1])
    with pytest.raises(x3d.X3DError):
        r.render_batch([(built[0], descs[0].camera, 1, 8, 1)])
    films, st, lat = r.render_batch([])
    assert films == [] and st.num_samples == 0


def test_batch_through_a_wavefront_form_renderer(x3d, batch):
    """A renderer of the wavefront form (the default for UNIDIRECT_MIS) batches the same films."""
    descs, built, items = batch
    plain, _, _ = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=6, seed=9).render_batch(items)
    wf, _, _ = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=6, seed=9).render_batch(items)
    for a, b in zip(plain, wf):
        assert np.array_equal(_bits(a), _bits(b))


def _generations(items, capacity):
    """Generation count of the wavefront batch: whole samples of the jobs' pixels packed, fewest
    samples first, into generations of max(capacity, largest film) paths."""
    cap = max([capacity] + [w * h for (_, _, w, h, _) in items])
    order = sorted(range(len(items)), key=lambda k: items[k][2] * items[k][3] * items[k][4])
    gens, used = 0, 0
    for k in order:
        npix, left = items[k][2] * items[k][3], items[k][4]
        while left:
            fit = (cap - used) // npix
            if fit == 0:
                gens, used = gens + 1, 0
                continue
            c = min(fit, left)
            used, left = used + c * npix, left - c
    return gens + (used > 0)


@pytest.mark.parametrize("kw", [dict(), dict(queue_capacity=5000), dict(queue_capacity=7000, workgroup_size=256),
                                dict(queue_capacity=1, workgroup_size=128, stack_depth=8)])
def test_wavefront_batch_films_identical_to_single_renders(x3d, batch, kw):
    descs, built, items = batch
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=6, seed=9, **kw)
    singles = [r.render(s, cam, w, h, spp)[0] for (s, cam, w, h, spp) in items]
    gens = _generations(items, kw.get("queue_capacity", 1 << 26))
    for rep in range(2):  # capture, then replay
        print(f"[gpu] wavefront batch {kw} pass {rep}: {gens} generations", flush=True)
        films, st, lat = r.render_batch(items)
        for k, (f, ref) in enumerate(zip(films, singles)):
            assert np.array_equal(_bits(f), _bits(ref)), (k, np.abs(f - ref).max())
        assert st.generations == gens and st.alone_jobs == 0 and st.tiles == 0
        assert st.num_samples == sum(w * h * spp for (_, _, w, h, spp) in items)
        assert st.closest_queries >= sum(w * h for (_, _, w, h, _) in items)
        assert all(0 < x <= st.kernel_ms * 1.01 + 0.05 for x in lat), (lat, st.kernel_ms)


def test_wavefront_batch_matches_oracle_and_autotunes(x3d, oracle_mod, batch):
    descs, built, items = batch
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=6, seed=3, queue_capacity=4000)
    films, st, _ = r.render_batch(items[:3] + items[-2:])
    assert st.generations > 1
    for k, (f, d) in enumerate(zip(films, descs[:3] + descs[-1:])):
        ref = oracle_mod.Oracle(d).render(oracle_mod.TR_UNIDIRECT_MIS, d.width, d.height, d.spp, seed=3, max_depth=6)
        assert np.array_equal(_bits(f), _bits(ref)), k
    rep = r.autotune_batch(items)
    assert rep.candidates == 6 and rep.best_ms > 0 and rep.workgroup_size in (64, 128, 256)
    again, _, _ = r.render_batch(items[:3] + items[-2:])
    for a, b in zip(films, again):
        assert np.array_equal(_bits(a), _bits(b))


def test_wavefront_batch_of_mixed_scenes(x3d):
    """One generation holding paths of scenes with one emitter, several emitters of unequal power
    and area, a sky light, a furnace and two triangle-heavy rooms: each job reads its own lights
    and materials per path, so every film equals the job rendered alone."""
    descs = [scenes.box_scene(), scenes.multi_light_scene(), scenes.sky_scene(), scenes.furnace_sphere(),
             scenes.indirect_scene(level=3), scenes.spheres_scene(count=12, target_tris=30000, seed=77)]
    built = [x3d.Scene(d) for d in descs]
    sizes = [(40, 24, 3), (33, 20, 4), (24, 24, 5), (16, 16, 2), (48, 32, 2), (31, 17, 3)]
    items = [(b, d.camera, w, h, spp) for b, d, (w, h, spp) in zip(built, descs, sizes)]
    for cap in (1 << 26, 2500):
        r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=7, seed=5, queue_capacity=cap)
        singles = [r.render(*it)[0] for it in items]
        films, st, _ = r.render_batch(items)
        print(f"[gpu] mixed-scene wavefront batch, capacity {cap}: {st.generations} generations", flush=True)
        assert st.generations == _generations(items, cap)
        for k, (f, ref) in enumerate(zip(films, singles)):
            assert np.array_equal(_bits(f), _bits(ref)), (cap, k)
            assert f[:, :3].max() > 0, k


def test_wavefront_batch_random_packings(x3d, batch):
    """Seeded random batches (film sizes 2-70 pixels a side, 1-9 spp, generation sizes from below
    the largest film up): slices that start a job mid-way through its samples, generations of one
    slice, jobs spread over many generations; every film equals its job rendered alone."""
    descs, built, _ = batch
    rng = np.random.default_rng(1234)
    for trial in range(4):
        n = int(rng.integers(3, 9))
        items = []
        for _ in range(n):
            k = int(rng.integers(0, len(built)))
            w, h, spp = int(rng.integers(2, 71)), int(rng.integers(2, 71)), int(rng.integers(1, 10))
            items.append((built[k], descs[k].camera, w, h, spp))
        cap = int(rng.choice([1, 97, 1000, 4096, 20000]))
        r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=4, seed=trial, queue_capacity=cap)
        singles = [r.render(*it)[0] for it in items]
        films, st, _ = r.render_batch(items)
        print(f"[gpu] random wavefront batch {trial}: {n} jobs, capacity {cap}, {st.generations} generations",
              flush=True)
        assert st.generations == _generations(items, cap)
        for k, (f, ref) in enumerate(zip(films, singles)):
            assert np.array_equal(_bits(f), _bits(ref)), (trial, k)
