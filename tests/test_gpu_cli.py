"""The command-line driver end to end on the GPU: a scene written as the
reference's files (PFM views and depth maps, CONTOUR camera files, a config
file), run through lfrv_cli, compared with the same sequence over the C-ABI
in this process and with the FP64 oracle.

IBR_gradient follows gradientIBR.cpp:12-150, IBR_color super-resolution.cpp:54-338.
"""
import subprocess
import sys

import numpy as np
import pytest

import lfrv
import oracle
from lfrv import io as lio
from lfrv import pipelines
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _write_cam(path, P):
    P = np.asarray(P, np.float64).reshape(3, 4)
    with open(path, "w") as f:
        f.write("CONTOUR\n")
        for r in range(3):
            f.write(" ".join(f"{v:.17g}" for v in P[r]) + "\n")


def _write_scene(d, s, s_rmv=-1):
    """Views and depths at indices 0..V (skipping s_rmv), cameras likewise,
    and the target camera at nviews (+1 when a view is removed), as
    import_warps_from_depth reads it (import.cpp:254-258)."""
    V = s["V"]
    idx = [i for i in range(V + (1 if s_rmv != -1 else 0)) if i != s_rmv]
    for v, i in enumerate(idx):
        lio.save_image(str(d / f"v_{i:02d}.pfm"), s["views"][v])
        lio.save_image(str(d / f"d_{i:02d}.pfm"), np.repeat(s["depths"][v][None], 3, 0))
        _write_cam(d / f"cam_{i:02d}.txt", s["P_views"][v])
    _write_cam(d / f"cam_{V + (1 if s_rmv != -1 else 0):02d}.txt", s["P_target"])
    return 0, idx[-1]


def _config(d, s, s_min, s_max, s_rmv, algorithm, extra):
    cfg = d / "scene.cfg"
    lines = [f"algorithm {algorithm}", f"lf_name {d}/v_%02d.pfm", f"depth_name {d}/d_%02d.pfm",
             f"cam_name {d}/cam_%02d.txt", f"tau_name {d}/tau_%02d.pfm", f"outdir {d}/out", "outfile u.png",
             f"s_min {s_min}", f"s_max {s_max}", f"s_rmv {s_rmv}", f"dsf {s['dsf']}", "nchannels 3",
             "sigma_sensor 1.0"] + extra
    cfg.write_text("\n".join(lines) + "\n")
    return cfg


def _run_cli(cfg):
    print(f"lfrv_cli -config {cfg}", flush=True)
    r = subprocess.run([lio.CLI_PATH, "-config", str(cfg)], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("s_rmv", [-1, 2])
def test_cli_ibr_gradient_matches_abi_and_oracle(tmp_path, small_scene, s_rmv):
    s = small_scene
    s_min, s_max = _write_scene(tmp_path, s, s_rmv)
    cfg = _config(tmp_path, s, s_min, s_max, s_rmv, "IBR_gradient",
                  ["dt_alpha 1.0", "dt_beta 0.5", "ibr_system normal_equations"])
    _run_cli(cfg)
    out = tmp_path / "out"
    U_cli = lio.load_image(str(out / "output.pfm"))
    # the same sequence over the C-ABI in this process
    prm = pipelines.gradient_ibr_params(s, dt_alpha=1.0, dt_beta=0.5, sigma_sensor=np.float32(1.0) / np.float32(255))
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.clear_solution()
        ctx.warps_from_depth()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        U0 = ctx.get_solution()
        ctx.init_u_gradient()
        st = ctx.cg_solve(1e-5, 5000)
        U = ctx.get_solution()
        tx = ctx.get_buffer("TAU_X").reshape(s["V"], s["h"], s["w"])
        ty = ctx.get_buffer("TAU_Y").reshape(s["V"], s["h"], s["w"])
    assert all(st["converged"])
    assert np.max(np.abs(U_cli - U)) <= 1e-6 * max(1.0, np.abs(U).max())
    # files of the driver: the PNG of the result and of the initial image, the tau PFMs
    want_png = np.clip(np.trunc(U.astype(np.float64) * 255.0), 0, 255) / 255.0
    assert np.max(np.abs(lio.load_image(str(out / "u.png")) - want_png)) <= 1.0 / 255.0 + 1e-7
    want0 = np.clip(np.trunc(U0.astype(np.float64) * 255.0), 0, 255) / 255.0
    assert np.max(np.abs(lio.load_image(str(out / "u_init_filled.png")) - want0)) <= 1.0 / 255.0 + 1e-7
    idx = [i for i in range(s_max + 1) if i != s_rmv]
    for v, i in enumerate(idx):
        t = lio.load_image(str(tmp_path / f"tau_{i:02d}.pfm"))
        assert np.array_equal(t[0], tx[v]) and np.array_equal(t[1], ty[v]) and np.array_equal(t[2], ty[v])
    # and the oracle's pipeline on the same inputs
    r = oracle.gibr_pipeline(s["views"], s["depths"], s["P_views"], s["P_target"], s["dsf"], 1 / 255.0, 1.0,
                             1.0, 0.5, 1e-5, 2000)
    assert rel_err(U_cli, r["U"]) < 1e-3


def test_cli_ibr_gradient_default_is_the_screened_poisson_system(tmp_path, small_scene):
    """A reference config without dt_beta (Config_data's default 0,
    config.cpp:61): possion_conj_grad solves the reference's own system
    (I + D^T D) u = u0 + D^T g (solver.cu:361-505, lambda = 1 at :541), u0 the
    hole-filled initial image and g the splatted luminance gradient, whatever
    dt_alpha / dt_beta say."""
    s = small_scene
    s_min, s_max = _write_scene(tmp_path, s)
    cfg = _config(tmp_path, s, s_min, s_max, -1, "IBR_gradient", [])
    _run_cli(cfg)
    U_cli = lio.load_image(str(tmp_path / "out" / "output.pfm"))
    H, W, V, h, w = s["H"], s["W"], s["V"], s["h"], s["w"]
    prm = pipelines.gradient_ibr_params(s, dt_alpha=1.0, dt_beta=0.0, sigma_sensor=np.float32(1.0) / np.float32(255))
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.clear_solution()
        ctx.warps_from_depth()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        U0 = ctx.get_solution().astype(np.float64)
        ctx.init_u_gradient()
        gx = ctx.get_buffer("U_GRAD_X").reshape(H, W)
        gy = ctx.get_buffer("U_GRAD_Y").reshape(H, W)
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        st = ctx.poisson_cg(1.0, 1e-5, 5000, precond=False)
        U = ctx.get_solution()
    assert all(st["converged"])
    assert np.max(np.abs(U_cli - U)) <= 1e-6 * max(1.0, np.abs(U).max())
    # the FP64 oracle's CG on (I + D^T D) from the same u0 and g
    wz = np.zeros(tx.shape)
    ob = oracle.rhs(W, H, s["dsf"], tx, ty, s["views"], wz, 0.0, 1.0, U0, 1.0, gx, gy)
    oX, ost = oracle.pcg(W, H, s["dsf"], tx, ty, wz, 0.0, 1.0, 1.0, ob, U0, 1e-8, 5000, False)
    assert all(ost["converged"])
    assert rel_err(U_cli, oX) < 1e-4
    # the gradient term is really in it: the result is not u0, and it is not
    # the normal-equations solution either
    assert rel_err(U_cli, U0) > 1e-3
    with lfrv.Context(pipelines.gradient_ibr_params(s, dt_alpha=1.0, dt_beta=0.5)) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.ibr_gradient(1e-5, 5000)
        assert rel_err(U_cli, ctx.get_solution()) > 1e-3


def test_cli_ibr_gradient_from_tau_files(tmp_path, small_scene):
    """Warps read from tau / dpart PFMs (import_warps) instead of depth maps."""
    s = small_scene
    with lfrv.Context(pipelines.gradient_ibr_params(s, dt_beta=0.5)) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.warps_from_depth()
        V, h, w = s["V"], s["h"], s["w"]
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        dpx = ctx.get_buffer("DPART_X").reshape(V, h, w)
        dpy = ctx.get_buffer("DPART_Y").reshape(V, h, w)
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        ctx.init_u_gradient()
        ctx.cg_solve(1e-5, 5000)
        U = ctx.get_solution()
    for v in range(V):
        lio.save_image(str(tmp_path / f"v_{v:02d}.pfm"), s["views"][v])
        lio.save_image(str(tmp_path / f"t_{v:02d}.pfm"), np.stack([tx[v], ty[v], ty[v]]))
        # sigma_z in R is not read by the solver; dtau/dz in G and B
        lio.save_image(str(tmp_path / f"p_{v:02d}.pfm"), np.stack([np.zeros_like(dpx[v]), dpx[v], dpy[v]]))
    cfg = tmp_path / "scene.cfg"
    cfg.write_text("\n".join([
        "algorithm IBR_gradient", f"lf_name {tmp_path}/v_%02d.pfm", f"tau_name {tmp_path}/t_%02d.pfm",
        f"dpart_name {tmp_path}/p_%02d.pfm", f"outdir {tmp_path}/out", "outfile u.pfm", "s_min 0",
        f"s_max {V - 1}", f"dsf {s['dsf']}", "nchannels 3", "sigma_sensor 1", "dt_beta 0.5",
        "ibr_system normal_equations"]) + "\n")
    _run_cli(cfg)
    U_cli = lio.load_image(str(tmp_path / "out" / "u.pfm"))
    assert np.max(np.abs(U_cli - U)) <= 1e-6 * max(1.0, np.abs(U).max())


def _sr_driver(ctx, W, H, niter, nmeta):
    """super-resolution.cpp:177-300 restated over the C-ABI in the reference's
    float (cuflt) arithmetic, with host copies of the previous solutions.
    Returns the solution and what the loop did: FISTA steps, re-weighting
    passes, how each pass's FISTA loop ended (0 niter steps, 1 small energy,
    2 small relative change, 3 few moving pixels), the last energy and
    whether a re-weighting rule ended the solve."""
    f = np.float32
    EPS, MAXD, PCT = f(1e-4), f(1e-3), f(0.01)
    prev_meta = ctx.get_solution()
    prev = prev_meta.copy()
    energy = prev_e = prev_e_rw = f(0)
    info = dict(steps=0, meta=0, stops=[], converged=0)
    for r in range(nmeta):
        info["meta"] = r + 1
        ctx.compute_weights()
        ctx.sr_init_regularizer_weight()
        energy = f(ctx.sr_primal_energy()[0])
        prev_e = energy
        stop = 0
        for k in range(niter):
            ctx.sr_iteration_fista()
            info["steps"] += 1
            energy = f(ctx.sr_primal_energy()[0])
            if energy < EPS:
                stop = 1
                break
            if abs((energy - prev_e) / energy) < EPS:
                stop = 2
                break
            prev_e = energy
            if f(ctx.sr_count_moving(prev, MAXD)) / f(W * H) < PCT:
                stop = 3
                break
            prev = ctx.get_solution()
        info["stops"].append(stop)
        if energy < EPS:
            if prev_e_rw < EPS:
                info["converged"] = 1
                break
            prev_e_rw = energy
            continue
        if f(ctx.sr_count_moving(prev_meta, MAXD)) / f(W * H) < PCT:
            info["converged"] = 1
            break
        prev_meta = ctx.get_solution()
    info["energy"] = energy
    return ctx.get_solution(), info


def test_cli_ibr_color_matches_abi(tmp_path):
    d = lfrv.synth_desc(1, 5)
    d.w = d.h = 40
    d.dsf = d.render_scale = 2
    d.focal = 40.0
    d.baseline = 0.7 * 4.0 / 80
    s = lfrv.synth(d)
    s_min, s_max = _write_scene(tmp_path, s)
    niter, nmeta = 15, 3
    cfg = _config(tmp_path, s, s_min, s_max, -1, "IBR_color",
                  ["lambda 0.05", "lambda_max_factor 100", f"niter {niter}", f"nmeta_iter {nmeta}",
                   "dt_alpha 1", "dt_beta 0.3", "gradient_step 1.0"])
    _run_cli(cfg)
    U_cli = lio.load_image(str(tmp_path / "out" / "output.pfm"))
    prm = pipelines.color_sr_params(s, lambda_=0.05, lambda_max_factor=100.0, inner_iterations=niter, dt_beta=0.3,
                                    gradient_step=1.0, sigma_sensor=np.float32(1.0) / np.float32(255))
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.clear_solution()
        pipelines.color_sr_setup(ctx)
        U, _ = _sr_driver(ctx, s["W"], s["H"], niter, nmeta)
    assert np.max(np.abs(U_cli - U)) <= 1e-5


def _color_scene(dsf, w=40, flat=False):
    d = lfrv.synth_desc(1, 5)
    d.w = d.h = w
    d.dsf = d.render_scale = dsf
    d.focal = float(w)
    d.baseline = 0.7 * 4.0 / (w * dsf)
    s = lfrv.synth(d)
    if flat:  # one colour everywhere: TV and gradient terms vanish
        s["views"][:] = np.array([0.2, 0.5, 0.7], np.float32)[None, :, None, None]
    return s


@pytest.mark.parametrize("dsf,inner,beta,niter,nmeta,flat,scale,stops", [
    (2, 3, 0.3, 15, 3, False, 1.0, {0}),      # niter reached in every pass; odd inner count
    (2, 2, 0.0, 120, 4, False, 1.0, {0, 2, 3}),  # long passes: the relative-change and moving rules
    (1, 1, 0.3, 120, 3, False, 1.0, {2, 3}),  # dsf = 1 data term
    (3, 5, 0.3, 60, 2, False, 1.0, {0}),
    (2, 1, 0.3, 0, 2, False, 1.0, {0}),       # no FISTA step: the passes only re-weight
    (2, 1, 0.3, 20, 3, True, 1e-9, {1}),      # flat views, tiny weights: the small-energy rules
])
def test_sr_solve_matches_the_driver_loop(dsf, inner, beta, niter, nmeta, flat, scale, stops):
    """lfrv_sr_solve (stopping rules on the device, steps enqueued ahead and
    skipped once a rule fired) against the same loop driven from the host,
    step by step, over the C-ABI: the same solution bit for bit, the same
    step count, stop reasons and energy.  `stops` are the ways the passes'
    FISTA loops end in each case (0 niter steps, 1 small energy, 2 small
    relative change, 3 few moving pixels), so every rule is exercised."""
    s = _color_scene(dsf, flat=flat)
    prm = pipelines.color_sr_params(s, lambda_=0.05 * scale, lambda_max_factor=100.0, inner_iterations=inner,
                                    dt_alpha=scale, dt_beta=beta * scale, gradient_step=1.0,
                                    sigma_sensor=np.float32(1.0) / np.float32(255))
    res = []
    for device_loop in (False, True):
        print(f"sr loop dsf={dsf} inner={inner} niter={niter} nmeta={nmeta} device={device_loop}", flush=True)
        with lfrv.Context(prm) as ctx:
            pipelines.load_scene(ctx, s)
            ctx.clear_solution()
            pipelines.color_sr_setup(ctx)
            if device_loop:
                st = ctx.sr_solve(niter, nmeta)
                res.append((ctx.get_solution(), st))
            else:
                res.append(_sr_driver(ctx, s["W"], s["H"], niter, nmeta))
    (U_h, info), (U_d, st) = res
    print(info, st)
    assert set(info["stops"]) == stops
    assert st["fista_steps"] == info["steps"]
    assert st["meta_iterations"] == info["meta"]
    assert st["last_stop"] == (info["stops"][-1] if info["stops"] else 0)
    assert st["converged"] == info["converged"]
    assert np.float32(st["energy"]) == info["energy"]
    assert np.array_equal(U_d, U_h)


def test_sr_views_in_one_context_are_independent():
    """lfrv_clear_solution starts an IBR_color view from a fresh workspace
    (duals, G, normalisation buffer zero, FISTA alpha 1): a view solved after
    another one in the same context equals the first bit for bit."""
    s = _color_scene(2)
    prm = pipelines.color_sr_params(s, lambda_=0.05, lambda_max_factor=100.0, inner_iterations=2, dt_beta=0.3,
                                    gradient_step=1.0)
    out = []
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        for _ in range(2):
            ctx.clear_solution()
            pipelines.color_sr_setup(ctx)
            st = ctx.sr_solve(40, 2)
            out.append((ctx.get_solution(), st["fista_steps"]))
            # a few more steps outside the loop change alpha and the duals
            ctx.sr_init_regularizer_weight()
            for _ in range(3):
                ctx.sr_iteration_fista()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
