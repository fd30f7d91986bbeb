"""The oracle's path-replay adjoint (gradient of a film loss with respect to material albedo and
GGX roughness) against central finite differences of the oracle's own FP64 renders with common
random numbers: the same PCG streams and the GGX sampling density held fixed, so every render is
a smooth function of the evaluated parameters and the FD quotient converges to the exact
derivative of the estimator. This pins the derivative formulas (d log f of both lobes, the
suffix sums of the path replay) without reference to the implementation."""

import numpy as np

from x3dcore2_amd import scenes


def _loss(o, w, h, spp, depth, target):
    img = o.render(16, w, h, spp, seed=5, max_depth=depth, prec=1)[:, :3]
    return ((img - target) ** 2).sum(), img


def test_adjoint_matches_central_differences(oracle_mod):
    w = h = 16
    spp, depth, step = 4, 3, 1e-3
    truth_a, truth_r = scenes.adjoint_truth()
    o = oracle_mod.Oracle(scenes.adjoint_scene())
    scenes.set_adjoint_params(o, truth_a, truth_r)
    target = o.render(16, w, h, spp, seed=5, max_depth=depth, prec=1)[:, :3]
    a0 = np.full((16, 3), 0.5, np.float32)
    r0 = np.full(16, 0.4, np.float32)
    scenes.set_adjoint_params(o, a0, r0)
    _, img = _loss(o, w, h, spp, depth, target)
    dl = np.zeros((w * h, 4), np.float32)
    dl[:, :3] = 2 * (img - target)
    ga, gr = scenes.tie_gradient(o.render_backward(w, h, spp, dl, seed=5, max_depth=depth, prec=1))
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
            scenes.set_adjoint_params(o, ap, rp)
            lp, _ = _loss(o, w, h, spp, depth, target)
            scenes.set_adjoint_params(o, am, rm)
            lm, _ = _loss(o, w, h, spp, depth, target)
            if c < 3:
                fa[m, c] = (lp - lm) / (2 * step)
            else:
                fr[m] = (lp - lm) / (2 * step)
    assert np.abs(fa).max() > 0.1 and np.abs(fr).max() > 0.1
    assert np.allclose(ga, fa, rtol=1e-3, atol=1e-4 * np.abs(fa).max()), np.c_[ga.ravel(), fa.ravel()]
    assert np.allclose(gr, fr, rtol=1e-3, atol=1e-4 * np.abs(fr).max()), np.c_[gr, fr]


def test_adjoint_zero_at_target_and_linear_in_dl(oracle_mod):
    """dl = 0 gives a zero gradient; the gradient is linear in dLoss/dfilm."""
    w = h = 12
    o = oracle_mod.Oracle(scenes.adjoint_scene())
    rng = np.random.default_rng(1)
    d1 = rng.normal(size=(w * h, 4)).astype(np.float32)
    d2 = rng.normal(size=(w * h, 4)).astype(np.float32)
    g1 = o.render_backward(w, h, 2, d1, seed=3, max_depth=3, prec=1)
    g2 = o.render_backward(w, h, 2, d2, seed=3, max_depth=3, prec=1)
    g12 = o.render_backward(w, h, 2, (d1 + 2 * d2).astype(np.float32), seed=3, max_depth=3, prec=1)
    assert not o.render_backward(w, h, 2, np.zeros_like(d1), seed=3, max_depth=3, prec=1).any()
    # (d1 + 2 d2 is rounded to float32 once)
    assert np.allclose(g12, g1 + 2 * g2, rtol=1e-5, atol=1e-6 * np.abs(g12).max())
    # mixtures and the emitter carry no parameters of their own
    assert not g1[2::3][:16].any() and not g1[3 * 16:].any()
