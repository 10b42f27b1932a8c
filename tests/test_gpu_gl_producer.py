"""The OpenGL warp producer on the GPU (lfrv_warps_from_target_depth) against
the oracle's restatement: the target depth map by depth splatting, tau with
its -1 / -2 markers, the deformation weights and dtau/dz; then the gradient
IBR solve on those warps against the oracle's pipeline on the oracle's warps."""
import numpy as np
import pytest

import lfrv
import oracle
from lfrv import pipelines
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _scene(w=48, h=40, seed=4):
    d = lfrv.synth_desc(2, seed)
    d.w, d.h = w, h
    d.focal = 1.5 * w      # disparities up to a few pixels, with occlusions between the planes
    return lfrv.synth(d)


def _grads(s, seed):
    r = np.random.default_rng(seed)
    return [(r.uniform(-0.02, 0.02, (s["h"], s["w"])).astype(np.float32),
             r.uniform(-0.02, 0.02, (s["h"], s["w"])).astype(np.float32)) for _ in range(s["V"])]


@pytest.mark.parametrize("with_grad,skip", [(False, -1), (True, -1), (False, 4)])
def test_gl_producer_matches_oracle(with_grad, skip):
    s = _scene()
    V, h, w = s["V"], s["h"], s["w"]
    W, H = s["W"], s["H"]
    g = _grads(s, 9) if with_grad else None
    print(f"gl producer grad={with_grad} skip={skip}", flush=True)
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        pipelines.load_scene(ctx, s)
        if g:
            for v in range(V):
                ctx.set_depth_gradient(v, *g[v])
        ctx.warps_from_target_depth(skip)
        ud = ctx.get_buffer("TARGET_DEPTH").reshape(H, W)
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        wd = ctx.get_buffer("TAU_WEIGHT").reshape(V, h, w)
        dpx = ctx.get_buffer("DPART_X").reshape(V, h, w)
    gx = np.stack([a for a, _ in g]) if g else None
    gy = np.stack([b for _, b in g]) if g else None
    oud = oracle.gl_target_depth(s["depths"], s["P_views"], s["P_target"], W, H, skip, gx, gy)
    # pixel centres on a quad edge or equal depths may go either way between FP32 and FP64
    diff = np.abs(ud - oud) > 1e-3 * np.maximum(1.0, oud)
    assert diff.mean() < 5e-3, diff.mean()
    assert (oud > 0).mean() > 0.9
    otx, oty, owd, odx, _ = oracle.gl_warps(s["depths"], s["P_views"], s["P_target"], ud, 10.0, gx, gy)
    same = (tx >= 0) == (otx >= 0)
    assert same.mean() > 0.998
    assert np.all(tx[~(otx >= 0) & same] == otx[~(otx >= 0) & same])   # the same marker, -1 or -2
    both = (tx >= 0) & (otx >= 0)
    assert np.max(np.abs(tx[both] - otx[both])) < 2e-3
    assert np.max(np.abs(ty[both] - oty[both])) < 2e-3
    assert np.max(np.abs(wd[both] - owd[both])) < 1e-3 * 30
    assert np.max(np.abs(dpx - odx)) < 1e-3 * max(1.0, np.abs(odx).max())
    assert (tx == -1).any() and (tx == -2).any()    # both markers occur in this scene


def test_gradient_ibr_on_gl_warps_matches_oracle():
    s = _scene(40, 32, seed=6)
    V, h, w = s["V"], s["h"], s["w"]
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.clear_solution()
        ctx.warps_from_target_depth(-1)
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        dpx = ctx.get_buffer("DPART_X").reshape(V, h, w)
        dpy = ctx.get_buffer("DPART_Y").reshape(V, h, w)
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        ctx.init_u_gradient()
        st = ctx.cg_solve(1e-5, 2000)
        U = ctx.get_solution()
    r = oracle.gibr_pipeline(s["views"], s["depths"], s["P_views"], s["P_target"], 1, 1 / 255.0, 1.0, 1.0, 0.5,
                             1e-5, 2000, taus=(tx, ty, dpx, dpy))
    assert all(st["converged"])
    assert rel_err(U, r["U"]) < 1e-4
    assert abs(oracle.psnr(U, s["target"]) - oracle.psnr(r["U"], s["target"])) < 0.01
