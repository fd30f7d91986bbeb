"""The oracle's firefly filter (sampling_task::run, renderer.cpp:46-83; FireFlyMinDiff = 1.5,
renderer.h:118) on hand-made sample images whose outcome follows from the definition:

- a flat image passes unchanged (sum = spp x value);
- an isolated bright interior sample is capped, component by component, at the mean squared
  norm of its 8 neighbours plus 1.5, in its own sample image only;
- a bright sample with a bright neighbour (difference below 1.5), one on the image edge, or one
  whose brightness sits two pixels from another bright sample outside its 3 x 3 block, behaves
  as the definition says.
The reference renderer's own filter sits in renderer.cpp, which cannot be built here (Qt), so
these closed-form cases pin the oracle; the GPU film is then compared with the oracle bit for bit.
"""

import numpy as np
import pytest


def _img(w, h, value):
    return np.tile(np.asarray(value, np.float32), (h, w, 1)).astype(np.float32)


def test_flat_image_unchanged(oracle_mod):
    orc = oracle_mod.Oracle()
    x = np.stack([_img(7, 5, (0.3, 0.2, 0.1))] * 4)
    out = orc.firefly_accumulate(x)
    ref = np.zeros((5, 7, 3), np.float32)
    for k in range(4):
        ref = ref + x[k]
    assert np.array_equal(out, ref)


def test_isolated_bright_sample_is_capped(oracle_mod):
    orc = oracle_mod.Oracle()
    bg = np.float32(0.1)
    x = np.stack([_img(6, 6, (bg, bg, bg)), _img(6, 6, (bg, bg, bg))])
    x[0, 2, 3] = (3.0, 0.05, 2.0)      # interior firefly in sample 0 only
    out = orc.firefly_accumulate(x)
    r = np.float32(bg * bg) + np.float32(bg * bg) + np.float32(bg * bg)   # neighbour squared norm
    cap = np.float32(np.float32(1.0 / 8.0) * np.float32(8 * r)) + np.float32(1.5)
    assert abs(cap - 1.53) < 1e-6
    want = np.array([min(cap, 3.0), 0.05, min(cap, 2.0)], np.float32) + bg
    assert np.allclose(out[2, 3], want, rtol=0, atol=1e-6)
    # every other pixel: plain sums
    mask = np.ones((6, 6), bool)
    mask[2, 3] = False
    assert np.array_equal(out[mask], (x[0] + x[1])[mask])


def test_non_fireflies_unchanged(oracle_mod):
    orc = oracle_mod.Oracle()
    x = _img(8, 6, (0.1, 0.1, 0.1))[None].repeat(1, 0)
    x[0, 2, 2] = (3.0, 0.0, 0.0)       # bright ...
    x[0, 2, 3] = (2.9, 0.0, 0.0)       # ... with a bright neighbour: 9 - 8.41 < 1.5
    x[0, 0, 5] = (4.0, 0.0, 0.0)       # on the top edge
    x[0, 4, 6] = (4.0, 0.0, 0.0)       # isolated: capped
    out = orc.firefly_accumulate(x)
    assert np.array_equal(out[2, 2], x[0, 2, 2]) and np.array_equal(out[2, 3], x[0, 2, 3])
    assert np.array_equal(out[0, 5], x[0, 0, 5])
    assert out[4, 6, 0] < 1.6 and out[4, 6, 0] > 1.5


@pytest.mark.parametrize("spp", [1, 3])
def test_filter_is_per_sample_image(oracle_mod, spp):
    orc = oracle_mod.Oracle()
    x = np.stack([_img(5, 5, (0.2, 0.2, 0.2)) for _ in range(spp)])
    x[-1, 2, 2] = (5.0, 5.0, 5.0)
    out = orc.firefly_accumulate(x)
    assert out[2, 2, 0] < 0.2 * (spp - 1) + 1.7
    if spp > 1:
        # the other sample images of that pixel enter unchanged
        assert out[2, 2, 0] > 0.2 * (spp - 1) + 1.5
