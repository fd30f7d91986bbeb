"""GPU film compositors (x3d_compositor: clamp_compositor / aces_compositor,
compositor.cpp:29-98) against the oracle's compositor byte for byte, with the same exposure
bits and pixel count:

- films of many shapes (1 x 1, ragged sizes around the 1024-pixel chunk, more chunks than the
  final tree's 256 lanes, 1920 x 1080) with a wide dynamic range;
- special values: NaN, +-inf, negatives, -0, zero pixels (not counted by auto exposure),
  alpha-only pixels, alpha outside [0, 1], a film with no counted pixel (NaN exposure);
- clamp, ACES with fixed exposure (including the default 0) and ACES with auto exposure, the
  host entry point and the device one;
- the rendered film of x3d_render_device composed in place on the device.
"""

import numpy as np
import pytest

from x3dcore2_amd import scenes
from x3dcore2_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (7, 3), (33, 31), (32, 32), (1025, 1), (600, 500), (1920, 1080)]
MODES = [("clamp", False, 1.0), ("aces", True, 0.0), ("aces", False, 0.7), ("aces", False, 0.0)]


def _bits(x):
    return np.float32(x).view(np.uint32)


def _film(w, h, seed):
    rng = np.random.default_rng(seed)
    f = rng.lognormal(-1.0, 2.0, (w * h, 4)).astype(np.float32)
    f[:, 3] = np.float32(1.0) / np.float32(1 + seed % 64)
    return f


def _special_film():
    w, h = 37, 29
    f = _film(w, h, 5)
    vals = np.array([np.nan, np.inf, -np.inf, -1.0, -0.0, 0.0, 1e-6, 2e-5, 1e20, 1e38], np.float32)
    rng = np.random.default_rng(6)
    idx = rng.choice(w * h, size=300, replace=False)
    f[idx[:100], rng.integers(0, 3, 100)] = rng.choice(vals, 100)
    f[idx[100:150]] = 0.0                       # not counted
    f[idx[150:200], :3] = 0.0                    # alpha only: counted
    f[idx[200:250], 3] = rng.choice(np.array([-0.5, 1.5, 300.0, np.nan], np.float32), 50)
    f[idx[250:300]] = [1e-5, -1e-5, 0.0, 1e-5]   # within 1e-5 of 0: not counted
    return f, w, h


def _compositor(x3d, kind, auto, e):
    c = x3d.Compositor(x3d.COMPOSITOR_CLAMP if kind == "clamp" else x3d.COMPOSITOR_ACES)
    c.enable_auto_exposure(auto)
    c.set_exposure(e)
    return c


def _check(oracle_mod, x3d, film, w, h, kind, auto, e):
    c = _compositor(x3d, kind, auto, e)
    frame, st = c.commit(film, w, h)
    ref, re, rn = oracle_mod.compose(film, w, h, oracle_mod.COMPOSE_CLAMP if kind == "clamp" else
                                     oracle_mod.COMPOSE_ACES, auto, e)
    assert np.array_equal(frame, ref), (np.argwhere(frame != ref)[:5], frame[frame != ref][:5])
    assert _bits(st.exposure) == _bits(re) and st.counted == rn
    return frame, st


@pytest.mark.parametrize("kind,auto,e", MODES)
@pytest.mark.parametrize("w,h", SHAPES)
def test_compose_matches_oracle(x3d, oracle_mod, w, h, kind, auto, e):
    print(f"[gpu] compose {kind} auto={auto} e={e} {w}x{h}", flush=True)
    _check(oracle_mod, x3d, _film(w, h, w + 7 * h), w, h, kind, auto, e)


@pytest.mark.parametrize("kind,auto,e", MODES)
def test_compose_special_values(x3d, oracle_mod, kind, auto, e):
    print(f"[gpu] compose special values {kind} auto={auto} e={e}", flush=True)
    film, w, h = _special_film()
    _check(oracle_mod, x3d, film, w, h, kind, auto, e)


def test_compose_at_gamma_thresholds(x3d, oracle_mod):
    """Every 8-bit gamma step: channel values within 64 ulps of (k/255)^2.2, k = 1 .. 255 (the
    GPU reads the byte from a threshold table; the oracle evaluates pow per value)."""
    g = np.float32(1.0) / np.float32(2.2)
    k = np.arange(1, 256)
    c = ((k / 255.0) ** (1.0 / np.float64(g))).astype(np.float32).view(np.uint32).astype(np.int64)
    u = (c[:, None] + np.arange(-64, 65)[None, :]).reshape(-1)
    x = np.minimum(u, 0x3F800000).astype(np.uint32).view(np.float32)
    n = x.size + (-x.size) % 3
    v = np.resize(x, n).reshape(-1, 3)
    film = np.concatenate([v, np.full((v.shape[0], 1), 0.5, np.float32)], 1)
    w = film.shape[0]
    _check(oracle_mod, x3d, film, w, 1, "clamp", False, 1.0)


def test_compose_nothing_counted(x3d, oracle_mod):
    film = np.zeros((64 * 20, 4), np.float32)
    frame, st = _check(oracle_mod, x3d, film, 64, 20, "aces", True, 0.0)
    assert st.counted == 0 and np.isnan(st.exposure) and not frame.any()


def test_compose_rendered_film_on_device(x3d, oracle_mod):
    desc = scenes.spheres_scene(count=12, target_tris=30000, seed=77)
    gs = x3d.Scene(desc)
    w, h, spp = 320, 240, 4
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=6, seed=9)
    film_d = DeviceArray(w * h * 16)
    frame_d = DeviceArray(w * h * 4)
    r.render_device(gs, desc.camera, w, h, spp, film_d.ptr)
    film = film_d.to_host(np.float32, (w * h, 4))
    c = x3d.Compositor(x3d.COMPOSITOR_ACES)
    for _ in range(2):  # the second call reuses the compositor's scratch buffers
        st = c.commit_device(film_d.ptr, w, h, frame_d.ptr)
        frame = frame_d.to_host(np.uint8, (w * h, 4))
        ref, re, rn = oracle_mod.compose(film, w, h, oracle_mod.COMPOSE_ACES)
        assert np.array_equal(frame, ref)
        assert _bits(st.exposure) == _bits(re) and st.counted == rn == w * h
    assert frame[:, :3].max() > 100 and (frame[:, 3] == 63).all()
    print(f"[gpu] rendered 320x240 film: exposure {st.exposure:.4f}, exposure pass {st.exposure_ms:.3f} ms, "
          f"compose {st.compose_ms:.3f} ms", flush=True)


def test_compose_errors(x3d):
    c = x3d.Compositor(x3d.COMPOSITOR_CLAMP)
    with pytest.raises(x3d.X3DError):
        c.commit(np.zeros((0, 4), np.float32), 0, 5)
    with pytest.raises(x3d.X3DError):
        x3d.Compositor(3)
