"""The structured disparity model of cocolib's vtv_sr (warp then box
downsample, vtv_sr_kernels.cu:72-135, 767-802; compiled into the reference but
not called by its unstructured light-field path) as the semantic cross-check of
the tau-footprint model that path uses (set_B_bilinear) for pure translations.

A view pixel X of the structured model averages dsf bilinear samples of u at
X dsf + j + s (integer-origin pixels, j < dsf); in corner-origin coordinates its
centre lands at tau = (X + 1/2) dsf + s.  With f the fractional part of the
first sample, the dsf samples put weights (1 - f, 1, ..., 1, f) / dsf on dsf + 1
consecutive pixels: exactly the footprint weights of set_B_bilinear at tau.  So
for pure translations the two models are one operator wherever all taps are
inside the image, for any image and any integer dsf."""
import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(7)


def _both(u, dsf, sx, sy):
    C, H, W = u.shape
    h, w = H // dsf, W // dsf
    X, Y = np.meshgrid(np.arange(w), np.arange(h))
    tx = ((X + 0.5) * dsf + sx).astype(np.float32)
    ty = ((Y + 0.5) * dsf + sy).astype(np.float32)
    foot = oracle.gather(W, H, dsf, tx, ty, u)
    ones = np.ones((H, W))
    struct = oracle.structured_warp_down(u, ones, sx, sy, ones, dsf)
    # interior: every tap of both models inside the image
    m = int(np.ceil(abs(sx))) + int(np.ceil(abs(sy))) + 2
    inner = np.zeros((h, w), bool)
    inner[(m + dsf) // dsf: h - (m + dsf) // dsf, (m + dsf) // dsf: w - (m + dsf) // dsf] = True
    return foot, struct, inner


@pytest.mark.parametrize("sx,sy", [(0.0, 0.0), (0.3, -0.7), (2.25, 1.5), (-1.8, 0.45)])
def test_dsf1_models_coincide(sx, sy):
    u = RNG.uniform(0, 1, (3, 24, 30))
    foot, struct, inner = _both(u, 1, sx, sy)
    assert inner.sum() > 100
    # tau reaches the footprint model in float32 (its ulp at 30 px is 2e-6 px)
    np.testing.assert_allclose(foot[:, inner], struct[:, inner], atol=5e-6)


@pytest.mark.parametrize("dsf", [2, 3, 4])
@pytest.mark.parametrize("sx,sy", [(0.0, 0.0), (0.37, -1.25), (-2.5, 0.8)])
def test_models_coincide_for_any_dsf(dsf, sx, sy):
    u = RNG.uniform(0, 1, (3, 12 * dsf, 16 * dsf))
    foot, struct, inner = _both(u, dsf, sx, sy)
    assert inner.sum() > 20
    np.testing.assert_allclose(foot[:, inner], struct[:, inner], atol=5e-6)


def test_models_part_where_the_structured_renormalises():
    """At the right border the structured model renormalises its bilinear
    weights while the footprint drops the outside taps and renormalises over
    the whole footprint: the models part there (the border is where the
    pinned equivalence stops)."""
    dsf, sx = 2, 0.6
    u = RNG.uniform(0, 1, (3, 16, 20))
    foot, struct, _ = _both(u, dsf, sx, 0.0)
    assert np.max(np.abs(foot[:, :, -1] - struct[:, :, -1])) > 1e-3
