"""Monte Carlo comparison helpers shared by the GPU image tests."""

import numpy as np


def gpu_replicates(x3d, desc, pt, w, h, spp, reps, scene=None, **kw):
    """`reps` independently seeded renders (of `scene`, else of a scene built from `desc`):
    rgb [reps, w*h, 3] float64."""
    scene = scene if scene is not None else x3d.Scene(desc)
    imgs = []
    for s in range(reps):
        r = x3d.Renderer(pt, seed=1000 + s, **kw)
        img, _ = r.render(scene, desc.camera, w, h, spp)
        imgs.append(img[:, :3].astype(np.float64))
    return np.stack(imgs)


def split_z(imgs, spp, other, other_spp):
    """z-scores of the GPU mean against another estimate of the same expectation. The first half
    of the replicates gives the mean, the second half (independent of it) the per-sample standard
    deviation, so that z has mean zero even for the skewed per-pixel distributions of these
    estimators; the other estimate (same estimator, other seeds) has the same per-sample
    variance. Returns z and the relative MSE of the GPU mean."""
    half = imgs.shape[0] // 2
    mean = imgs[:half].mean(0)
    sd = imgs[half:].std(0, ddof=1) * np.sqrt(spp)
    se = sd * np.sqrt(1.0 / (half * spp) + 1.0 / other_spp)
    ok = se > 1e-9
    z = (mean - other)[ok] / se[ok]
    return z, ((mean - other) ** 2)[ok].sum() / max((other ** 2)[ok].sum(), 1e-30)
