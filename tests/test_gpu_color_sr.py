"""GPU parity of the IBR_color path (TV primal-dual with warped data terms)
against the FP64 oracle, step by step through sr_synthesize_view."""
import numpy as np
import pytest

import lfrv
import oracle
from lfrv import pipelines
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _scene(dsf=2, kind=1, w=48):
    d = lfrv.synth_desc(1, 5)
    d.w = d.h = w
    d.dsf = dsf
    d.render_scale = dsf
    d.focal = float(w)
    d.baseline = 0.7 * 4.0 / (w * dsf)
    if kind == 2:
        d.kind = 2
        d.baseline = 2.0 * 4.0 / (w * dsf)
    return lfrv.synth(d)


@pytest.mark.parametrize("dsf,beta,gw,inner", [(2, 0.0, 0, 10), (2, 0.5, 1, 10), (3, 0.3, 0, 10), (1, 0.5, 1, 10),
                                              (2, 0.5, 0, 1), (4, 0.3, 0, 1), (4, 0.0, 1, 2), (3, 0.3, 0, 0)])
def test_sr_steps_match_oracle(dsf, beta, gw, inner):
    """inner = 1: for dsf >= 2 the data term and the only primal-dual iteration
    (with the over-relaxation) run as one kernel; inner > 1 fuses the first
    iteration; inner = 0 leaves the data term alone."""
    _steps_match_oracle(_scene(dsf=dsf, kind=2 if dsf == 1 else 1), beta, gw, inner)


def _c3_scene(w=96, seed=7):
    """The 5x5-view generator of the 512^2 -> 2048^2 workload (8 tilted planes,
    disparity -2..+2 px per grid step through the principal-point shift,
    views box-averaged x4) at w x w views: the focal length scales with w and
    f * baseline stays 10, so the disparities in view pixels are the same."""
    d = lfrv.synth_desc(3, seed)
    d.w = d.h = w
    d.focal = float(w)
    d.baseline = 10.0 / w
    return lfrv.synth(d)


@pytest.mark.parametrize("beta,inner", [(1.0, 1), (0.3, 2)])
def test_sr_steps_match_oracle_25_views_dsf4(beta, inner):
    """25 views and dsf 4 (the 512^2 -> 2048^2 configuration's generator, at
    96^2 -> 384^2): the pull kernels' per-window caps and the assembled
    anchor-grid operator at 25 views."""
    s = _c3_scene()
    assert s["V"] == 25 and s["dsf"] == 4
    _steps_match_oracle(s, beta, 0, inner, metas=2, steps=3)


def test_sr_full_size_128_to_256_matches_oracle():
    """The 3x3 128^2 -> 256^2 configuration at its full size: 20 FISTA steps of
    the benchmark's parameters (one primal-dual iteration, lambda 0.05,
    lambda_max_factor 100) against the oracle, U within 5e-4 and the energy
    within 1e-4 relative at the end."""
    s = lfrv.synth(1, seed=2024)
    _steps_match_oracle(s, 0.0, 0, 1, metas=1, steps=20, u_tol=5e-4, dual_tol=5e-4)


def _steps_match_oracle(s, beta, gw, inner, metas=2, steps=3, u_tol=2e-4, dual_tol=2e-4):
    dsf = s["dsf"]
    print(f"IBR_color V={s['V']} {s['w']}x{s['h']} dsf={dsf} beta={beta} gw={gw} inner={inner}", flush=True)
    V, h, w, H, W = s["V"], s["h"], s["w"], s["H"], s["W"]
    sig, lam, lmf = 1.0 / 255.0, 0.05, 100.0
    prm = pipelines.color_sr_params(s, lambda_=lam, lambda_max_factor=lmf, inner_iterations=inner, dt_beta=beta,
                                    gw_type=gw, gradient_step=1.0)
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.warps_from_depth()
        ctx.prepare()
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        dpx = ctx.get_buffer("DPART_X").reshape(V, h, w)
        dpy = ctx.get_buffer("DPART_Y").reshape(V, h, w)
        # oracle state mirrors coco_vtv_workspace
        st = oracle.SrState(W, H, np.zeros((3, H, W)))
        vis0 = np.zeros((H, W), np.uint8)
        ctx.compute_weights()
        om, gx, gy = oracle.sr_weights(W, H, dsf, tx, ty, dpx, dpy, st.U, vis0, st.PX[0], st.PY[0], sig, 1.0)
        st.PX[0], st.PY[0] = gx, gy
        assert rel_err(ctx.get_buffer("OMEGA").reshape(V, h, w), om) < 1e-6
        ctx.setup_visibility()
        cov = oracle.coverage(W, H, dsf, tx, ty)
        vis = (cov >= 1.0).astype(np.uint8)
        st.norm_mask += cov
        gvis = ctx.get_buffer("VISIBILITY").reshape(H, W)
        near = np.abs(cov - 1.0) < 1e-5
        assert np.array_equal(gvis[~near].astype(np.uint8), vis[~near])
        vis = np.where(near, gvis.astype(np.uint8), vis)
        ctx.compute_initial_image()
        U0, nm = oracle.initial_image(W, H, dsf, tx, ty, s["views"], om)
        st.norm_mask = nm
        U0 = oracle.push_pull(U0, vis)
        st.U = U0.copy()
        st.Uq = U0.copy()
        assert np.max(np.abs(ctx.get_solution() - U0)) < 1e-5
        for meta in range(metas):
            ctx.compute_weights()
            om, gx, gy = oracle.sr_weights(W, H, dsf, tx, ty, dpx, dpy, st.U, vis, st.PX[0], st.PY[0], sig, 1.0)
            st.PX[0], st.PY[0] = gx, gy
            assert rel_err(ctx.get_buffer("OMEGA").reshape(V, h, w), om) < 1e-5
            ctx.sr_init_regularizer_weight()
            g, st.norm_mask = oracle.regularizer_weight(W, H, dsf, tx, ty, om, sig, lmf * lam, lam, st.norm_mask)
            assert np.max(np.abs(ctx.get_buffer("REG_G").reshape(H, W) - g)) < 1e-5 * lmf * lam
            assert rel_err(ctx.get_buffer("NORM_MASK").reshape(H, W), st.norm_mask) < 1e-5
            e, parts = ctx.sr_primal_energy()
            oe, oparts = oracle.sr_energy(W, H, dsf, tx, ty, s["views"], om, sig, lam, 1.0, beta, gw, vis, st.U)
            assert abs(e - oe) / oe < 1e-4
            for k in range(steps):
                ctx.sr_iteration_fista()
                oracle.fista_step(W, H, dsf, tx, ty, s["views"], om, sig, 1.0, beta, 1.0, gw, lam, inner, g, st)
            U = ctx.get_solution()
            assert np.max(np.abs(U - st.U)) < u_tol, np.max(np.abs(U - st.U))
            assert np.max(np.abs(ctx.get_buffer("UQ").reshape(3, H, W) - st.Uq)) < 2 * u_tol
            assert np.max(np.abs(ctx.get_buffer("PX").reshape(3, H, W) - st.PX)) < dual_tol
            assert np.max(np.abs(ctx.get_buffer("PY").reshape(3, H, W) - st.PY)) < dual_tol
            assert rel_err(ctx.get_buffer("NORM_MASK").reshape(H, W), st.norm_mask) < 1e-5
        # the energy of the final iterate
        e, _ = ctx.sr_primal_energy()
        oe, _ = oracle.sr_energy(W, H, dsf, tx, ty, s["views"], om, sig, lam, 1.0, beta, gw, vis, st.U)
        assert abs(e - oe) / oe < 1e-4, (e, oe)


def test_sr_count_moving():
    s = _scene()
    prm = pipelines.color_sr_params(s)
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        pipelines.color_sr_setup(ctx)
        U = ctx.get_solution()
        prev = U.copy()
        rng = np.random.default_rng(0)
        idx = rng.choice(s["H"] * s["W"], 100, replace=False)
        prev.reshape(3, -1)[1, idx] += 2e-3
        prev.reshape(3, -1)[0, idx[:10]] += 5e-4
        assert ctx.sr_count_moving(prev, 1e-3) == 100


@pytest.mark.parametrize("dsf,inner,beta", [(2, 1, 0.3), (3, 1, 0.0), (1, 3, 0.3)])
def test_sr_iterate_equals_single_steps(dsf, inner, beta):
    """lfrv_sr_iterate(k) gives the solution, duals, over-relaxed solution and
    normalisation buffer of k lfrv_sr_iteration_fista calls, bit for bit (a
    fixed run writes the last two only in its last step)."""
    s = _scene(dsf=dsf, kind=2 if dsf == 1 else 1)
    prm = pipelines.color_sr_params(s, inner_iterations=inner, dt_beta=beta, gradient_step=1.0)
    out = []
    for batched in (False, True):
        with lfrv.Context(prm) as ctx:
            pipelines.load_scene(ctx, s)
            ctx.clear_solution()
            pipelines.color_sr_setup(ctx)
            ctx.sr_init_regularizer_weight()
            if batched:
                ctx.sr_iterate(7)
            else:
                for _ in range(7):
                    ctx.sr_iteration_fista()
            out.append((ctx.get_solution(), ctx.get_buffer("PX"), ctx.get_buffer("PY"), ctx.get_buffer("UQ"),
                        ctx.get_buffer("NORM_MASK")))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    assert np.abs(out[0][3].ravel() - out[0][0].ravel()).max() > 0  # the over-relaxed solution is not the solution
