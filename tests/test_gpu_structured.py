"""The GPU's footprint gather (lfrv_op_gather) against cocolib's structured
disparity model (warp then box downsample, vtv_sr_kernels.cu:72-135, 767-802)
for pure translations, where the two models are one operator inside the image."""
import numpy as np
import pytest

import lfrv
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dsf,sx,sy", [(1, 0.3, -0.7), (2, 1.25, 0.4), (3, -0.8, 2.1), (4, 0.37, -1.25)])
def test_gpu_gather_is_the_structured_warp(dsf, sx, sy):
    rng = np.random.default_rng(dsf)
    h, w = 14, 18
    H, W = h * dsf, w * dsf
    u = rng.uniform(0, 1, (3, H, W)).astype(np.float32)
    X, Y = np.meshgrid(np.arange(w), np.arange(h))
    tx = ((X + 0.5) * dsf + sx).astype(np.float32)
    ty = ((Y + 0.5) * dsf + sy).astype(np.float32)
    D = lfrv.DeviceArray
    dtx, dty, du, dv = D(tx), D(ty), D(u), D((3, h, w))
    assert lfrv.lib().lfrv_op_gather(W, H, w, h, dsf, dtx.ptr, dty.ptr, du.ptr, 3, dv.ptr, None) == 0
    lfrv.device_sync()
    v = dv.numpy()
    ones = np.ones((H, W))
    want = oracle.structured_warp_down(u, ones, sx, sy, ones, dsf)
    m = (int(np.ceil(abs(sx))) + int(np.ceil(abs(sy))) + 2 + dsf) // dsf
    inner = (slice(None), slice(m, h - m), slice(m, w - m))
    assert np.max(np.abs(v[inner] - want[inner])) < 1e-5
