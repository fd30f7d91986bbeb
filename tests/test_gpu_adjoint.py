"""Path-replay adjoint on the GPU (x3d_render_backward through the C ABI):

- against the oracle's adjoint on the same PCG streams: the FP32 path (same operations as the
  GPU) and the FP64 one;
- against central finite differences (h = 1e-3) of the GPU's own forward renders with common
  random numbers and the GGX sampling density held fixed, on the inverse-rendering workload
  (256 x 256, 32 spp, depth 4, 16 surfaces x (albedo rgb, roughness)): within 1e-3 relative;
- reproducible bit for bit from run to run;
- 200 Adam steps from a grey, medium-rough start recover the seeded parameters (tightly for
  every parameter the image is sensitive to).
"""

import numpy as np
import pytest

from x3dcore2_amd import scenes

pytestmark = pytest.mark.gpu


def _setup(x3d, w, h):
    truth_a, truth_r = scenes.adjoint_truth()
    desc = scenes.adjoint_scene(width=w, height=h)
    sc = x3d.Scene(desc)
    return desc, sc, truth_a, truth_r


def test_adjoint_matches_oracle(x3d, oracle_mod):
    w, h, spp, depth = 32, 32, 8, 4
    desc, sc, ta, tr = _setup(x3d, w, h)
    rng = np.random.default_rng(2)
    dl = rng.normal(size=(w * h, 4)).astype(np.float32)
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=depth, seed=7)
    orc = oracle_mod.Oracle(desc)
    for params in ((None, None), (ta, tr)):
        if params[0] is not None:
            scenes.set_adjoint_params(sc, *params)
            scenes.set_adjoint_params(orc, *params)
        g, st = r.backward(sc, desc.camera, w, h, spp, dl)
        g32 = orc.render_backward(w, h, spp, dl, seed=7, max_depth=depth, prec=0)
        g64 = orc.render_backward(w, h, spp, dl, seed=7, max_depth=depth, prec=1)
        scale = np.abs(g64).max()
        print(f"[gpu] adjoint vs oracle: max |gpu - fp32 path| {np.abs(g - g32).max() / scale:.2e}, "
              f"max |gpu - fp64| {np.abs(g - g64).max() / scale:.2e} (relative to max |grad|)", flush=True)
        assert st.num_samples == w * h * spp
        assert np.abs(g - g32).max() <= 1e-4 * scale
        # a renderer of the wavefront form (the default) runs the same replay
        gw, _ = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, form=x3d.FORM_WAVEFRONT, max_depth=depth, seed=7).backward(
            sc, desc.camera, w, h, spp, dl)
        assert np.array_equal(gw.view(np.uint32), g.view(np.uint32))
        assert np.abs(g - g64).max() <= 2e-3 * scale


def _loss(r, sc, desc, target, w, h, spp):
    img, _ = r.forward(sc, desc.camera, w, h, spp)
    d = img[:, :3].astype(np.float64) - target
    return (d * d).sum(), img


def test_adjoint_matches_central_fd_and_is_reproducible(x3d):
    w, h, spp, depth, step = 256, 256, 32, 4, 1e-3
    desc, sc, ta, tr = _setup(x3d, w, h)
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=depth, seed=11)
    scenes.set_adjoint_params(sc, ta, tr)
    target, _ = r.forward(sc, desc.camera, w, h, spp)
    target = target[:, :3].astype(np.float64)
    a0 = np.full((16, 3), 0.5, np.float32)
    r0 = np.full(16, 0.4, np.float32)
    scenes.set_adjoint_params(sc, a0, r0)
    _, img = _loss(r, sc, desc, target, w, h, spp)
    dl = np.zeros((w * h, 4), np.float32)
    dl[:, :3] = 2 * (img[:, :3] - target)
    g, st = r.backward(sc, desc.camera, w, h, spp, dl)
    g2, _ = r.backward(sc, desc.camera, w, h, spp, dl)
    assert np.array_equal(g.view(np.uint32), g2.view(np.uint32))
    print(f"[gpu] backward 256x256x32: {st.kernel_ms:.2f} ms", flush=True)
    ga, gr = scenes.tie_gradient(g)
    fa, fr = np.zeros((16, 3)), np.zeros(16)
    for m in range(16):
        for c in range(4):
            ap, rp, am, rm = a0.copy(), r0.copy(), a0.copy(), r0.copy()
            if c < 3:
                ap[m, c] += step
                am[m, c] -= step
            else:
                rp[m] += step
                rm[m] -= step
            scenes.set_adjoint_params(sc, ap, rp)
            lp, _ = _loss(r, sc, desc, target, w, h, spp)
            scenes.set_adjoint_params(sc, am, rm)
            lm, _ = _loss(r, sc, desc, target, w, h, spp)
            if c < 3:
                fa[m, c] = (lp - lm) / (2 * step)
            else:
                fr[m] = (lp - lm) / (2 * step)
    scenes.set_adjoint_params(sc, a0, r0)
    ea = np.abs(ga - fa) / np.maximum(np.abs(fa), 1e-12)
    er = np.abs(gr - fr) / np.maximum(np.abs(fr), 1e-12)
    big_a, big_r = np.abs(fa) > 1e-2 * np.abs(fa).max(), np.abs(fr) > 1e-2 * np.abs(fr).max()
    print(f"[gpu] adjoint vs FD: albedo max rel err {ea[big_a].max():.2e} over {big_a.sum()} params, "
          f"roughness {er[big_r].max():.2e} over {big_r.sum()}", flush=True)
    assert ea[big_a].max() < 1e-3 and er[big_r].max() < 1e-3
    # the rest agree in absolute terms
    assert np.abs(ga - fa).max() < 1e-4 * np.abs(fa).max() + 1e-3 * np.abs(fa).max()
    assert np.abs(gr - fr).max() < 1e-4 * np.abs(fr).max() + 1e-3 * np.abs(fr).max()


def test_adam_recovers_parameters(x3d):
    """200 Adam steps (lr 0.02, betas 0.8 / 0.9) from grey albedo 0.5 and roughness 0.4. The
    target is rendered from the seeded truth with the same seed and the same (start) GGX
    sampling density, so the loss is zero exactly at the truth. Parameters whose effect on the
    image is at least 5% of the strongest of their kind (film change per unit parameter change,
    measured at the truth) must be recovered tightly; faces that the camera barely sees (the
    blocks' back faces) only loosely."""
    w, h, spp, depth = 256, 256, 32, 4
    desc, sc, ta, tr = _setup(x3d, w, h)
    r = x3d.Renderer(x3d.PT_UNIDIRECT_MIS, max_depth=depth, seed=13)
    scenes.set_adjoint_params(sc, ta, tr)
    target, _ = r.forward(sc, desc.camera, w, h, spp)
    target = target[:, :3].astype(np.float64)
    truth = np.concatenate([ta.ravel(), tr]).astype(np.float64)
    sens = np.zeros(64)
    for k in range(64):
        q = truth.copy()
        q[k] += 1e-2
        scenes.set_adjoint_params(sc, q[:48].reshape(16, 3), q[48:])
        img, _ = r.forward(sc, desc.camera, w, h, spp)
        sens[k] = np.sqrt(((img[:, :3] - target) ** 2).sum()) / 1e-2
    ident = np.concatenate([sens[:48] >= 0.05 * sens[:48].max(), sens[48:] >= 0.05 * sens[48:].max()])
    p = np.concatenate([np.full(48, 0.5), np.full(16, 0.4)])
    m, v = np.zeros_like(p), np.zeros_like(p)
    lr, b1, b2 = 0.02, 0.8, 0.9
    losses = []
    for it in range(1, 201):
        scenes.set_adjoint_params(sc, p[:48].reshape(16, 3), p[48:])
        img, _ = r.forward(sc, desc.camera, w, h, spp)
        d = img[:, :3].astype(np.float64) - target
        losses.append((d * d).sum())
        dl = np.zeros((w * h, 4), np.float32)
        dl[:, :3] = 2 * d
        g, _ = r.backward(sc, desc.camera, w, h, spp, dl)
        ga, gr = scenes.tie_gradient(g)
        grad = np.concatenate([ga.ravel(), gr])
        m = b1 * m + (1 - b1) * grad
        v = b2 * v + (1 - b2) * grad * grad
        p -= lr * (m / (1 - b1 ** it)) / (np.sqrt(v / (1 - b2 ** it)) + 1e-8)
        p[:48] = np.clip(p[:48], 0.01, 0.99)
        p[48:] = np.clip(p[48:], 0.05, 1.0)
    err = np.abs(p - truth)
    ea, er = err[:48], err[48:]
    print(f"[gpu] Adam: loss {losses[0]:.3e} -> {losses[-1]:.3e}; {ident[:48].sum()} albedo / {ident[48:].sum()} "
          f"roughness parameters identifiable: |err| max {ea[ident[:48]].max():.4f} / {er[ident[48:]].max():.4f}; "
          f"all parameters |err| mean {ea.mean():.4f} / {er.mean():.4f}, max {ea.max():.4f} / {er.max():.4f}",
          flush=True)
    assert losses[-1] < 1e-3 * losses[0]
    assert ident[:48].sum() >= 20 and ident[48:].sum() >= 6
    assert ea[ident[:48]].max() < 0.01 and er[ident[48:]].max() < 0.02
    assert ea.mean() < 0.01 and er.mean() < 0.02
