"""Ragged and degenerate inputs on the GPU against the FP64 oracle: sizes that
are no multiple of any tile, a single view, tiny images, and a view whose depth
is invalid everywhere (import.cpp:292-296 sends all its pixels out of the
target), for both models."""
import numpy as np
import pytest

import lfrv
import oracle
from lfrv import pipelines
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _planes(w, h, dsf, grid, seed, dead_view=None):
    d = lfrv.synth_desc(2, seed)
    d.w, d.h, d.dsf, d.grid = w, h, dsf, grid
    d.focal = float(w)          # disparities of w/200 .. w/40 px per grid step
    d.render_scale = dsf
    s = lfrv.synth(d)
    if dead_view is not None:
        s["depths"][dead_view][:] = 0.0
    return s


GRAD_CASES = [
    (37, 23, 1, 3, None),   # 2x2 footprints: the assembled stencil with partial 64x8 tiles
    (29, 17, 2, 2, None),   # four views, entry-walk operator
    (21, 19, 3, 1, None),   # one view
    (6, 5, 1, 2, None),     # smaller than one tile
    (40, 24, 1, 3, 4),      # the centre view's depth invalid everywhere
]


@pytest.mark.parametrize("w,h,dsf,grid,dead", GRAD_CASES)
def test_gradient_ibr_ragged_inputs(w, h, dsf, grid, dead):
    s = _planes(w, h, dsf, grid, seed=11, dead_view=dead)
    print(f"IBR_gradient w={w} h={h} dsf={dsf} views={s['V']} dead={dead}", flush=True)
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        pipelines.load_scene(ctx, s)
        st = ctx.ibr_gradient(1e-5, 2000)
        U = ctx.get_solution()
        tx = ctx.get_buffer("TAU_X").reshape(s["V"], h, w)
    r = oracle.gibr_pipeline(s["views"], s["depths"], s["P_views"], s["P_target"], dsf, 1 / 255.0, 1.0,
                             1.0, 0.5, 1e-5, 2000)
    if dead is not None:
        assert np.all(tx[dead] < 0) and np.all(r["tx"][dead] < 0)
    assert all(st["converged"])
    assert np.all(np.isfinite(U))
    assert rel_err(U, r["U"]) < 1e-3
    assert abs(oracle.psnr(U, s["target"]) - oracle.psnr(r["U"], s["target"])) < 0.01


COLOR_CASES = [
    (29, 17, 2, 1),   # the data-term / primal-dual pass with partial 32x16 tiles
    (13, 11, 4, 1),   # dsf 4 on a target of 52 x 44
    (7, 5, 3, 2),     # smaller than one tile, two inner iterations
    (23, 19, 1, 3),   # dsf 1: entry pull + separate primal-dual steps
]


@pytest.mark.parametrize("w,h,dsf,inner", COLOR_CASES)
def test_color_sr_ragged_inputs(w, h, dsf, inner):
    s = _planes(w, h, dsf, 3, seed=7)
    V, H, W = s["V"], s["H"], s["W"]
    print(f"IBR_color w={w} h={h} dsf={dsf} inner={inner}", flush=True)
    sig, lam, lmf, beta = 1.0 / 255.0, 0.05, 100.0, 0.5
    prm = pipelines.color_sr_params(s, lambda_=lam, lambda_max_factor=lmf, inner_iterations=inner, dt_beta=beta,
                                    gradient_step=1.0)
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.clear_solution()
        pipelines.color_sr_setup(ctx)
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        dpx = ctx.get_buffer("DPART_X").reshape(V, h, w)
        dpy = ctx.get_buffer("DPART_Y").reshape(V, h, w)
        gvis = ctx.get_buffer("VISIBILITY").reshape(H, W).astype(np.uint8)
        # the oracle's state from the GPU's warps, as sr_synthesize_view sets it up
        st = oracle.SrState(W, H, np.zeros((3, H, W)))
        om, gx, gy = oracle.sr_weights(W, H, dsf, tx, ty, dpx, dpy, st.U, np.zeros((H, W), np.uint8), st.PX[0],
                                       st.PY[0], sig, 1.0)
        st.PX[0], st.PY[0] = gx, gy
        cov = oracle.coverage(W, H, dsf, tx, ty)
        vis = (cov >= 1.0).astype(np.uint8)
        near = np.abs(cov - 1.0) < 1e-5
        assert np.array_equal(gvis[~near], vis[~near])
        vis = np.where(near, gvis, vis)
        st.norm_mask += cov
        U0, st.norm_mask = oracle.initial_image(W, H, dsf, tx, ty, s["views"], om)
        st.U = oracle.push_pull(U0, vis)
        st.Uq = st.U.copy()
        assert np.max(np.abs(ctx.get_solution() - st.U)) < 1e-5
        ctx.sr_init_regularizer_weight()
        g, st.norm_mask = oracle.regularizer_weight(W, H, dsf, tx, ty, om, sig, lmf * lam, lam, st.norm_mask)
        ctx.sr_iterate(4)
        for _ in range(4):
            oracle.fista_step(W, H, dsf, tx, ty, s["views"], om, sig, 1.0, beta, 1.0, 0, lam, inner, g, st)
        U = ctx.get_solution()
        assert np.all(np.isfinite(U))
        assert np.max(np.abs(U - st.U)) < 2e-4
        assert np.max(np.abs(ctx.get_buffer("PX").reshape(3, H, W) - st.PX)) < 2e-4
        e, _ = ctx.sr_primal_energy()
        oe, _ = oracle.sr_energy(W, H, dsf, tx, ty, s["views"], om, sig, lam, 1.0, beta, 0, vis, st.U)
        assert abs(e - oe) <= 1e-4 * max(oe, 1e-12)
