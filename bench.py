"""Benchmark of the MI355X light-field solver on the headline configuration.

Workload (one step = one novel view): 9 synthetic views at 1024^2 RGB with
per-view depth of a seeded 8-tilted-plane scene, pinhole cameras on a 3x3
grid (baseline 0.05, f = 1000 px) and a random interior target camera.  The
step runs, from device-resident views / depths / cameras to the
device-resident solution: warps from depth, splat structure, visibility,
weights, weighted initial image with Poisson hole filling, target gradient,
and the Jacobi-PCG solve of the gradient-constraint normal equations to
relative residual 1e-5 (FP32 vectors, FP64 reductions).

Prints one JSON line.  --gpus above 1 is not supported (one GPU per box).
"""
import argparse
import json
import os
import platform
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "light-field-remote-vision_amd"))

import lfrv  # noqa: E402
from lfrv import pipelines  # noqa: E402

METRIC = "ms per novel view to rel. residual 1e-5 (9 views, 1024²); achieved HBM GB/s"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published peak
# one launch in TIMING_STRIDE of the roofline kernel carries dispatch-stamped
# HIP events in the timed region: a stamped dispatch waits for its completion
# signal, which costs several microseconds per launch on short kernels
TIMING_STRIDE = 8
TOL = 1e-5
# rocprofv3 summaries of the roofline kernels; each records the HBM bytes per
# launch measured with FETCH_SIZE / WRITE_SIZE (copy-calibrated)
PROFILE = os.path.join(ROOT, "profiles", "k_cgcg_st2_gradient_ibr_9views_1024.md")
PROFILE_C3 = os.path.join(ROOT, "profiles", "k_sr_step_st_color_sr_25views_2048.md")
PROFILE_C1 = os.path.join(ROOT, "profiles", "k_sr_step_st_color_sr_9views_256.md")
# the IBR_color workloads whose step kernel has a PMC record: workload -> (profile, kernel)
COLOR_TRAFFIC = {"color_sr_c3": (PROFILE_C3, "k_sr_step_st<5, true>"),
                 "color_sr_c1": (PROFILE_C1, "k_sr_step_st<3, false>")}


def measured_traffic(path=PROFILE):
    try:
        m = re.search(r"^traffic_bytes_per_launch: (\d+)", open(path).read(), re.M)
        return int(m.group(1)) if m else None
    except OSError:
        return None


def warp_producer_times(ctx, scene, reps=5):
    """Device time of the two depth -> tau producers on the bench scene, after
    the timed region: the import path (lfrv_warps_from_depth, the one the solve
    uses) and the OpenGL path (lfrv_warps_from_target_depth: target depth by
    quad splatting, then tau with its markers).  Algorithmic bytes per source
    pixel: the import path reads the depth (4 B) and writes tau_x, tau_y,
    dpart_x, dpart_y (16 B); the OpenGL path also writes the deformation
    weight (4 B) and, per target pixel, the target depth (4 B)."""
    n = scene["V"] * scene["w"] * scene["h"]
    N = scene["W"] * scene["H"]
    out = {}
    for name, run, nbytes in (("import_warps_from_depth", ctx.warps_from_depth, 20 * n),
                              ("opengl_warps_from_target_depth", lambda: ctx.warps_from_target_depth(-1),
                               24 * n + 4 * N)):
        run()
        ctx.sync()
        ctx.set_timing(True)
        for _ in range(reps):
            run()
        ctx.sync()
        ms, _ = ctx.get_timing("warp")
        ctx.set_timing(False)
        ms /= reps
        out[name] = dict(ms=round(ms, 4), bytes=nbytes, achieved_gbs=round(nbytes / (ms * 1e6), 1))
    return out


def hole_filling_times(scene, reps=3):
    """The Poisson hole filling (gradientIBR.cu:87-152) on the bench scene with
    a 256 x 256 block of invalid depth in every view: lfrv_compute_initial_image
    timed with and without the holes, after the timed region."""
    out = {}
    for name, holed in (("without_holes", False), ("with_holes", True)):
        s = dict(scene)
        if holed:
            d = scene["depths"].copy()
            d[:, 384:640, 384:640] = 0.0   # depth < 1e-3: invalid (import.cpp:292-296)
            s["depths"] = d
        ctx = lfrv.Context(pipelines.gradient_ibr_params(s))
        pipelines.load_scene(ctx, s)
        ms = []
        for r in range(reps + 1):
            ctx.clear_solution()
            ctx.warps_from_depth()
            ctx.prepare()
            ctx.setup_visibility()
            ctx.compute_weights()
            ctx.sync()
            t = time.perf_counter()
            ctx.compute_initial_image()
            ctx.sync()
            if r:
                ms.append(1e3 * (time.perf_counter() - t))
        holes, sweeps = ctx.hole_stats()
        ctx.close()
        out[name] = dict(initial_image_ms=round(float(np.median(ms)), 3), holes=holes, sweeps=sweeps)
    return out


def cell_valid_count(tx, ty, W, H):
    ok = (tx >= 0) & (ty >= 0) & (tx <= W) & (ty <= H)
    px = np.floor(tx - np.float32(0.5))
    py = np.floor(ty - np.float32(0.5))
    ok &= (px >= 0) & (py >= 0) & (px <= W - 1) & (py <= H - 1)
    return int(ok.sum())


def cpu_model():
    """CPU model name of the host (/proc/cpuinfo), else the platform's."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_threads():
    """Threads the oracle may use here: the box's share (OMP_NUM_THREADS when
    set, else the cores this process may run on), at most 16."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, min(16, n))


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def cpu_baseline(scene, threads):
    """The FP64 oracle on the host, measured: the whole novel view (its own
    warps from depth, visibility, weights, initial image with hole filling,
    target gradient, Jacobi-PCG to relative residual 1e-5) on all `threads`
    cores; then the same view's set-up plus 2 PCG iterations on one core."""
    oracle = _oracle()
    s = scene
    args = (s["views"], s["depths"], s["P_views"], s["P_target"], s["dsf"], 1.0 / 255.0, 1.0, 1.0, 0.5, TOL, 5000)
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    r = oracle.gibr_pipeline(*args)
    t_all = time.perf_counter() - t0
    # one core: the set-up and 2 PCG iterations of the same view
    oracle.set_threads(1)
    t0 = time.perf_counter()
    r1 = oracle.gibr_pipeline(*args[:-1], 2, taus=None)
    t_one = time.perf_counter() - t0
    oracle.set_threads(threads)
    iters = sum(r["stats"]["iterations"])
    one_core = dict(ms=round(1e3 * t_one, 1), pcg_iterations=sum(r1["stats"]["iterations"]),
                    sample="set-up of the view plus 2 PCG iterations per channel, FP64 oracle on 1 core")
    return dict(value=round(1e3 * t_all, 1), unit="ms per novel view", cores=threads, kind="port",
                sample=(f"the whole view measured (not scaled): FP64 oracle from depth maps and cameras to the "
                        f"PCG solution at rel. residual 1e-5 ({iters} iterations over 3 channels), OpenMP over "
                        f"rows on {threads} threads"),
                cpu=cpu_model(), nproc=os.cpu_count(), one_core=one_core, pcg_iterations=r["stats"]["iterations"])


def error_vs_oracle(scene, gpu, threads):
    """The GPU solution against the FP64 oracle run on the GPU's own warps
    (the same update equations on the same tau / dtau fields): relative
    solution error, relative energy error (each energy of its own solution)
    and PSNR difference to the ground truth."""
    oracle = _oracle()
    s = scene
    oracle.set_threads(threads)
    r = oracle.gibr_pipeline(s["views"], s["depths"], s["P_views"], s["P_target"], s["dsf"], 1.0 / 255.0, 1.0, 1.0,
                             0.5, TOL, 5000, taus=gpu["taus"])
    U = gpu["U"].astype(np.float64)
    rel = float(np.linalg.norm(U - r["U"]) / np.linalg.norm(r["U"]))
    e_rel = abs(gpu["energy"] - r["energy"]) / abs(r["energy"])
    p_orc = oracle.psnr(r["U"], s["target"])
    return dict(rel_err_vs_oracle=float("%.3g" % rel), rel_energy_err=float("%.3g" % e_rel),
                dpsnr_vs_oracle_db=round(gpu["psnr"] - p_orc, 5), oracle_psnr_db=round(p_orc, 3),
                oracle_energy=r["energy"], gpu_energy=gpu["energy"],
                oracle_pcg_iterations=r["stats"]["iterations"])


# IBR_color workloads: (synthetic scene preset, FISTA steps per view, dt_beta, description)
COLOR_WORKLOADS = {
    "color_sr_c3": (3, 500, 1.0, "5x5 views 512² -> 2048²"),
    "color_sr_c1": (1, 300, 0.0, "3x3 views 128² -> 256²"),
    "color_sr_c1_irls": (1, 300, 0.0, "3x3 views 128² -> 256²"),
}
# re-weighting passes of the workloads that run the driver's loop with its
# stopping rules (lfrv_sr_solve) instead of a fixed step count
IRLS_PASSES = {"color_sr_c1_irls": 3}


def host_driven_sr_loop(ctx, W, H, niter, nmeta):
    """The loop of super-resolution.cpp:177-300 driven from the host over the
    C-ABI the way the reference drives it: after each FISTA step the energy
    and the moving-pixel count come back to the host (the previous solution
    a host copy, as the reference's R_prev).  Returns the FISTA steps run."""
    f = np.float32
    EPS, MAXD, PCT = f(1e-4), f(1e-3), f(0.01)
    prev_meta = ctx.get_solution()
    prev = prev_meta.copy()
    prev_e_rw = f(0)
    steps = 0
    for _ in range(nmeta):
        ctx.compute_weights()
        ctx.sr_init_regularizer_weight()
        energy = prev_e = f(ctx.sr_primal_energy()[0])
        for _ in range(niter):
            ctx.sr_iteration_fista()
            steps += 1
            energy = f(ctx.sr_primal_energy()[0])
            if energy < EPS or abs((energy - prev_e) / energy) < EPS:
                break
            prev_e = energy
            if f(ctx.sr_count_moving(prev, MAXD)) / f(W * H) < PCT:
                break
            prev = ctx.get_solution()
        if energy < EPS:
            if prev_e_rw < EPS:
                break
            prev_e_rw = energy
            continue
        if f(ctx.sr_count_moving(prev_meta, MAXD)) / f(W * H) < PCT:
            break
        prev_meta = ctx.get_solution()
    return steps


def oracle_sr_view(scene, prm, steps, taus=None, budget_s=None, vis_ties=None):
    """One IBR_color view by the FP64 oracle: warps (its own, or `taus` =
    (tx, ty, dpx, dpy)), weights, visibility, push-pull initial image, TV
    weight, then FISTA steps.  With `budget_s` it stops after the first two
    steps when the rest would not fit in that many seconds.  `vis_ties` (the
    GPU's visibility) decides the pixels whose coverage is 1 to within 1e-5:
    the rule coverage >= 1 is a tie there, which FP32 and FP64 sums of the
    same footprint weights break either way."""
    oracle = _oracle()
    s = scene
    dsf, W, H = s["dsf"], s["W"], s["H"]
    sig = float(prm.sigma_sensor)
    t0 = time.perf_counter()
    if taus is None:
        tx, ty, dpx, dpy, _ = oracle.warps_from_depth(s["depths"], s["P_views"], s["P_target"], dsf)
    else:
        tx, ty, dpx, dpy = taus
    st = oracle.SrState(W, H, np.zeros((3, H, W)))
    om, gx, gy = oracle.sr_weights(W, H, dsf, tx, ty, dpx, dpy, st.U, np.zeros((H, W), np.uint8), st.PX[0],
                                   st.PY[0], sig, 1.0)
    st.PX[0], st.PY[0] = gx, gy
    cov = oracle.coverage(W, H, dsf, tx, ty)
    vis = (cov >= 1.0).astype(np.uint8)
    ties = np.abs(cov - 1.0) < 1e-5
    if vis_ties is not None:
        vis = np.where(ties, vis_ties.astype(np.uint8), vis)
    st.norm_mask += cov
    U0, st.norm_mask = oracle.initial_image(W, H, dsf, tx, ty, s["views"], om)
    st.U = oracle.push_pull(U0, vis)
    st.Uq = st.U.copy()
    g, st.norm_mask = oracle.regularizer_weight(W, H, dsf, tx, ty, om, sig, prm.lambda_ * prm.lambda_max_factor,
                                                prm.lambda_, st.norm_mask)
    t_setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    done = 0
    while done < steps:
        oracle.fista_step(W, H, dsf, tx, ty, s["views"], om, sig, prm.dt_alpha, prm.dt_beta, prm.gradient_step,
                          prm.gw_type, prm.lambda_, prm.inner_iterations, g, st)
        done += 1
        if budget_s is not None and done == 2:
            per = (time.perf_counter() - t1) / 2
            if t_setup + per * steps > budget_s:
                break
    t_steps = time.perf_counter() - t1
    e, _ = oracle.sr_energy(W, H, dsf, tx, ty, s["views"], om, sig, prm.lambda_, prm.dt_alpha, prm.dt_beta,
                            prm.gw_type, vis, st.U)
    return dict(U=st.U, energy=e, steps=done, t_setup=t_setup, t_steps=t_steps, ties=int(ties.sum()))


def cpu_baseline_sr(scene, prm, steps, threads, budget_s=30.0):
    """The FP64 oracle on the host for an IBR_color view (its own warps): the
    whole view measured when it fits in `budget_s`, else its set-up and two
    FISTA steps, scaled to the `steps` of the GPU run (said in `sample`)."""
    oracle = _oracle()
    oracle.set_threads(threads)
    r = oracle_sr_view(scene, prm, steps, budget_s=budget_s)
    if r["steps"] == steps:
        ms = 1e3 * (r["t_setup"] + r["t_steps"])
        sample = (f"the whole view measured (not scaled): FP64 oracle set-up ({r['t_setup']:.1f} s) and all "
                  f"{steps} FISTA steps, OpenMP on {threads} threads")
    else:
        per = r["t_steps"] / r["steps"]
        ms = 1e3 * (r["t_setup"] + per * steps)
        sample = (f"FP64 oracle set-up of one view ({r['t_setup']:.1f} s, measured) plus {r['steps']} FISTA steps "
                  f"({per * 1e3:.0f} ms each, measured), scaled to the {steps} steps of the GPU run; OpenMP on "
                  f"{threads} threads")
    return dict(value=round(ms, 1), unit="ms per novel view", cores=threads, kind="port", sample=sample,
                cpu=cpu_model(), nproc=os.cpu_count())


def error_vs_oracle_sr(scene, prm, taus, steps, U_gpu, e_gpu, psnr_gpu, threads, vis_gpu):
    """The GPU view after `steps` FISTA steps against the FP64 oracle's, on
    the GPU's warps (and its visibility where coverage ties with 1)."""
    oracle = _oracle()
    oracle.set_threads(threads)
    r = oracle_sr_view(scene, prm, steps, taus=taus, vis_ties=vis_gpu)
    U = U_gpu.astype(np.float64)
    p_orc = oracle.psnr(r["U"], scene["target"])
    return dict(steps_compared=steps,
                rel_err_vs_oracle=float("%.3g" % (np.linalg.norm(U - r["U"]) / np.linalg.norm(r["U"]))),
                max_abs_err_vs_oracle=float("%.3g" % np.max(np.abs(U - r["U"]))),
                rel_energy_err=float("%.3g" % (abs(e_gpu - r["energy"]) / abs(r["energy"]))),
                dpsnr_vs_oracle_db=round(psnr_gpu - p_orc, 5), oracle_psnr_db=round(p_orc, 3),
                visibility_ties=r["ties"])


def calib_copy(args):
    """A plain copy of --calib-copy MiB with k_copy16 (lfrv_dev_copy_kernel):
    the reference kernel of the PMC byte calibration."""
    if args.calib_copy > 0:
        nf = args.calib_copy * (1 << 20) // 4
        src = lfrv.DeviceArray((nf,))
        dst = lfrv.DeviceArray((nf,))
        lfrv.copy_kernel(dst, src)
        lfrv.device_sync()
        src.free()
        dst.free()


def bench_color_sr(args, workload):
    """One step = one IBR_color novel view: from device-resident views, depths
    and cameras, warps from depth, splat structure, weights, visibility,
    push-pull initial image and TV weight, then the FISTA / primal-dual steps
    (the data term, one primal-dual iteration and the over-relaxation each;
    for dsf >= 2 the data term is assembled once into anchor-grid couplings
    and each step is one stencil pass doing all three)."""
    preset, steps, beta, desc = COLOR_WORKLOADS[workload]
    nmeta = IRLS_PASSES.get(workload, 0)
    scene = lfrv.synth(preset, seed=args.seed)
    prm = pipelines.color_sr_params(scene, lambda_=0.05, lambda_max_factor=100.0, inner_iterations=1,
                                    dt_beta=beta, gradient_step=1.0)
    ctx = lfrv.Context(prm)
    pipelines.load_scene(ctx, scene)
    ctx.sync()
    irls = []

    def one_view():
        ctx.clear_solution()
        pipelines.color_sr_setup(ctx)
        if nmeta:
            irls.append(ctx.sr_solve(steps, nmeta))
        else:
            ctx.sr_init_regularizer_weight()
            ctx.sr_iterate(steps)
        ctx.sync()

    for _ in range(args.warmup):
        one_view()
    # the driver loop enqueues steps ahead of its stopping rules and the
    # device skips those behind a rule that fired: its kernels are timed in a
    # fixed-step pass below, where every launch does the full work
    ctx.set_timing(0 if nmeta else TIMING_STRIDE)
    per_view = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        one_view()
        per_view.append(1e3 * (time.perf_counter() - ts))
    t1 = time.perf_counter()
    ms_per_view = 1e3 * (t1 - t0) / args.steps
    U = ctx.get_solution()
    e_binned, ntiles = ctx.bin_stats()
    psnr = lfrv.psnr(U, scene["target"])
    if nmeta:
        ctx.set_timing(True)
        ctx.sr_init_regularizer_weight()
        for _ in range(50):
            ctx.sr_iteration_fista()
        ctx.sync()
    dt_ms, dt_n = ctx.get_timing("dataterm")
    pd_ms, pd_n = ctx.get_timing("pd_step")
    ctx.set_timing(False)
    producers = warp_producer_times(ctx, scene)
    convergence = None
    if not nmeta:
        # convergence of the fixed-step solve, after the timed region: the
        # primal energy (FP64) and the PSNR to the ground truth at checkpoints
        ctx.clear_solution()
        pipelines.color_sr_setup(ctx)
        ctx.sr_init_regularizer_weight()
        e0, _ = ctx.sr_primal_energy()
        traj = [dict(step=0, energy=round(e0, 6), psnr_db=round(lfrv.psnr(ctx.get_solution(), scene["target"]), 3))]
        done = 0
        for k in sorted({max(1, steps // 10), steps // 4, steps // 2, steps}):
            ctx.sr_iterate(k - done)
            done = k
            ek, _ = ctx.sr_primal_energy()
            traj.append(dict(step=k, energy=round(ek, 6),
                             psnr_db=round(lfrv.psnr(ctx.get_solution(), scene["target"]), 3)))
        emin = min(t["energy"] for t in traj)
        convergence = dict(trajectory=traj, rel_energy_decrease=round((e0 - traj[-1]["energy"]) / e0, 6),
                           final_over_min_energy=round(traj[-1]["energy"] / emin, 6),
                           same_as_timed=bool(np.array_equal(ctx.get_solution(), U)))
    host_loop = None
    if nmeta:
        # the same view with the loop driven from the host, for comparison
        ctx.clear_solution()
        pipelines.color_sr_setup(ctx)
        ctx.sync()
        th = time.perf_counter()
        host_steps = host_driven_sr_loop(ctx, scene["W"], scene["H"], steps, nmeta)
        ctx.sync()
        host_loop = dict(ms_per_view=round(1e3 * (time.perf_counter() - th), 3), fista_steps=host_steps,
                         same_result=bool(np.array_equal(ctx.get_solution(), U)))
    ctx.close()
    calib_copy(args)
    N = scene["W"] * scene["H"]
    e = e_binned
    ks = scene["dsf"] + 1
    if ks == 2:
        # data-term pull, per cell-valid entry: e_src, e_tx, e_ty, e_w (16 B) and the
        # source view (12 B; + its Laplacian, 12 B, with the gradient term); per
        # target pixel: U, coverage read, G, Uq, norm_mask written (44 B)
        dt_kernel = "k_sr_dataterm<2>"
        dt_bytes = (28 + (12 if beta > 0 else 0)) * e + 44 * N
    else:
        # one pass per outer step: the assembled data term and the primal-dual
        # iteration with the over-relaxation.  Per anchor (W + dsf) x (H + dsf)
        # the couplings G9 (5 planes) and bI (3), with the gradient term G9' (5,
        # unweighted) and one right-hand side b (3) for both terms instead of bI;
        # per target pixel U, coverage, p (6 planes), g
        # read (44 B), p and the new U written (36 B).  The over-relaxed Uq and
        # norm_mask (16 B more) are written by the run's last step only, which the
        # sampled timing (launches 0, 8, 16, ... counted over all views) never
        # stamps at 300 or 500 steps per view (the last step's index is 3 or 7 mod 8)
        NA = (scene["W"] + ks - 1) * (scene["H"] + ks - 1)
        dt_kernel = f"k_sr_step_st<{ks}, {'true' if beta > 0 else 'false'}>"  # <dsf + 1, gradient term>
        dt_bytes = 4 * NA * (8 + (5 if beta > 0 else 0)) + 80 * N
    # fused PD step: Uq, px, py, g, F read; px, py, new primal written (88 B
    # per pixel), with the outer step's over-relaxation fused in: U read, and
    # U and the over-relaxed Uq written (+24 B)
    pd_bytes = 112 * N
    dt_s = dt_ms / max(dt_n, 1) / 1e3
    pd_s = pd_ms / max(pd_n, 1) / 1e3
    kern = {
        dt_kernel: dict(bytes_per_launch=dt_bytes, avg_launch_us=round(dt_s * 1e6, 2),
                              achieved_gbs=round(dt_bytes / dt_s / 1e9, 1), launches=dt_n),
    }
    if pd_n:  # separate primal-dual launches (dsf = 1; dsf >= 2 runs them inside the step pass)
        kern["k_pd_step"] = dict(bytes_per_launch=pd_bytes, avg_launch_us=round(pd_s * 1e6, 2),
                                 achieved_gbs=round(pd_bytes / pd_s / 1e9, 1), launches=pd_n)
    dom = dt_kernel if dt_ms >= pd_ms else "k_pd_step"
    k = kern[dom]
    if nmeta:
        last = irls[-1]
        metric = (f"ms per novel view ({desc}, up to {nmeta} re-weighting passes of up to {steps} FISTA steps, "
                  f"stopping rules on the device); achieved HBM GB/s")
        workload_desc = (f"IBR_color driver loop (super-resolution.cpp:177-300) on the device: {desc} RGB, up to "
                         f"{nmeta} passes x {steps} FISTA steps of one primal-dual iteration, energy and "
                         f"moving-pixel rules after every step, lambda 0.05, lambda_max_factor 100, dt_beta {beta}")
    else:
        metric = f"ms per novel view ({desc}, {steps} TV primal-dual steps); achieved HBM GB/s"
        workload_desc = (f"IBR_color TV-L2 super-resolution with warped data terms: {desc} RGB, "
                         f"{steps} FISTA steps of one primal-dual iteration, lambda 0.05, "
                         f"lambda_max_factor 100, dt_beta {beta}")
    line = {
        "metric": metric,
        "value": round(ms_per_view, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_view, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload_desc,
                   "views": scene["V"], "width": scene["W"], "height": scene["H"], "dsf": scene["dsf"],
                   "seed": args.seed},
        "roofline": {"bound": "hbm", "kernel": dom,
                     "achieved": k["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(k["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": (measured_traffic(COLOR_TRAFFIC[workload][0])
                                 if COLOR_TRAFFIC.get(workload, (None, None))[1] == dom else None),
                     "bytes_per_launch": k["bytes_per_launch"], "avg_launch_us": k["avg_launch_us"],
                     "timed_launches": k["launches"], "timing_stride": TIMING_STRIDE},
        "ms_per_step_median": round(float(np.median(per_view)), 3),
        "ms_per_step_p10": round(float(np.percentile(per_view, 10)), 3),
        "ms_per_step_p90": round(float(np.percentile(per_view, 90)), 3),
        "psnr_vs_ground_truth_db": round(psnr, 3),
        "kernels": dict(kern, binned_entries=e_binned, tiles=ntiles),
        "warp_producers": producers,
    }
    if convergence:
        line["convergence"] = convergence
    if nmeta:
        line["irls"] = dict(meta_iterations=last["meta_iterations"], fista_steps=last["fista_steps"],
                            last_stop=last["last_stop"], converged=last["converged"],
                            energy=round(float(last["energy"]), 6), host_driven_loop=host_loop)
    if not args.no_cpu_baseline:
        threads = host_threads()
        if not nmeta:
            # error against the oracle: the whole view when the oracle runs it in
            # seconds, else the first steps of a view run again on the GPU
            cmp_steps = steps if scene["W"] * scene["H"] * scene["V"] * steps <= 9 * 256 * 256 * 300 else 2
            ctx = lfrv.Context(prm)
            pipelines.load_scene(ctx, scene)
            ctx.clear_solution()
            pipelines.color_sr_setup(ctx)
            ctx.sr_init_regularizer_weight()
            ctx.sr_iterate(cmp_steps)
            Ug = ctx.get_solution()
            eg, _ = ctx.sr_primal_energy()
            shp = (scene["V"], scene["h"], scene["w"])
            taus = tuple(ctx.get_buffer(b).reshape(shp) for b in ("TAU_X", "TAU_Y", "DPART_X", "DPART_Y"))
            vis_g = ctx.get_buffer("VISIBILITY").reshape(scene["H"], scene["W"])
            ctx.close()
            line.update(error_vs_oracle_sr(scene, prm, taus, cmp_steps, Ug, eg, lfrv.psnr(Ug, scene["target"]),
                                           threads, vis_g))
        line["cpu_baseline"] = cpu_baseline_sr(scene, prm, irls[-1]["fista_steps"] if nmeta else steps, threads)
    print(json.dumps(line))


# ---- batched serving: many uneven solves at once -------------------------------

def c4_problems(seed, n):
    """n solves with seeded uneven sizes and view counts: target side uniform
    in [256, 1536] (a multiple of 16), 2x2 .. 5x5 views (4 .. 25), half
    gradient-domain PCG solves (dsf 1, to rel. residual 1e-5) and half
    IBR_color solves (dsf 2, 100 FISTA steps), all scenes from the tilted-plane
    generator."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(rng.integers(256, 1537)) // 16 * 16
        grid = int(rng.integers(2, 6))
        kind = "cg" if rng.random() < 0.5 else "pd"
        out.append(dict(kind=kind, side=side, grid=grid, seed=int(seed * 1000 + i)))
    return out


def c4_desc(pr):
    if pr["kind"] == "cg":
        d = lfrv.synth_desc(2, pr["seed"])
        d.grid = pr["grid"]
        d.w = d.h = pr["side"]
        d.focal = 1000.0 * pr["side"] / 1024.0
        return d
    w = pr["side"] // 2
    d = lfrv.synth_desc(3, pr["seed"])
    d.grid = pr["grid"]
    d.w = d.h = w
    d.dsf = 2
    d.render_scale = 2
    d.focal = float(w)
    d.baseline = 10.0 / w
    return d


PD_STEPS = 100


def c4_contexts(problems, cfg):
    """One context per solve with its scene's views, depths and cameras in
    device memory (rendered on the GPU)."""
    from lfrv.warmstart import SceneOnDevice
    ctxs = []
    for pr in problems:
        d = c4_desc(pr)
        sc = SceneOnDevice(d)
        scene = dict(V=sc.V, w=d.w, h=d.h, dsf=d.dsf)
        kw = dict(tile_w=cfg["tile_w"], tile_h=cfg["tile_h"], pull_lds_kb=cfg["pull_lds_kb"])
        if pr["kind"] == "cg":
            prm = pipelines.gradient_ibr_params(scene, **kw)
        else:
            prm = pipelines.color_sr_params(scene, lambda_=0.05, lambda_max_factor=100.0, inner_iterations=1,
                                            dt_beta=0.3, gradient_step=1.0, **kw)
        ctx = lfrv.Context(prm)
        sc.load(ctx)
        sc.free()
        ctxs.append(ctx)
    return ctxs


def c4_serial(problems, ctxs):
    """The plain form: the solves one after another, each finished before the next."""
    lat = []
    sols = []
    t0 = time.perf_counter()
    for pr, ctx in zip(problems, ctxs):
        if pr["kind"] == "cg":
            ctx.ibr_gradient(TOL, 5000)
        else:
            ctx.clear_solution()
            pipelines.color_sr_setup(ctx)
            ctx.sr_init_regularizer_weight()
            ctx.sr_iterate(PD_STEPS)
        ctx.sync()
        lat.append(1e3 * (time.perf_counter() - t0))
    total = time.perf_counter() - t0
    for ctx in ctxs:
        sols.append(ctx.get_solution())
    return total, lat, sols


def c4_batched(problems, ctxs, chunk):
    from lfrv.batch import Batch
    b = Batch(chunk=chunk, use_graph=True)
    for pr, ctx in zip(problems, ctxs):
        if pr["kind"] == "cg":
            b.add_cg(ctx, TOL, 5000)
        else:
            b.add_pd(ctx, PD_STEPS)
    return b


def bench_batch(args):
    """One step = one batch of 64 uneven solves, from device-resident views,
    depths and cameras to the device-resident solutions."""
    from lfrv import autotune
    n = 64
    problems = c4_problems(args.seed, n)
    # autotuner on a calibration batch of 8 of the solves
    calib = problems[:8]

    def run_calib(cfg):
        ctxs = c4_contexts(calib, cfg)
        b = c4_batched(calib, ctxs, cfg["chunk"])
        b.run()
        lfrv.device_sync()
        t = time.perf_counter()
        b.run()
        lfrv.device_sync()
        dt = time.perf_counter() - t
        b.close()
        for c in ctxs:
            c.close()
        return dt

    sig = "batch_c4:" + ",".join(f"{p['kind']}{p['side']}x{p['grid']}" for p in calib)
    t_tune = time.perf_counter()
    cfg, trials = autotune.tune(run_calib, sig, reps=2, use_cache=not args.no_tune_cache,
                                log=lambda m: print(m, file=sys.stderr, flush=True))
    t_tune = time.perf_counter() - t_tune
    print(f"autotune: {cfg} in {t_tune:.1f} s ({len(trials)} trials)", file=sys.stderr, flush=True)
    ctxs = c4_contexts(problems, cfg)
    lfrv.device_sync()
    # the plain form: one solve after another (warm-up pass first)
    c4_serial(problems, ctxs)
    t_serial, lat_serial, sols_serial = c4_serial(problems, ctxs)
    # the batched form
    b = c4_batched(problems, ctxs, cfg["chunk"])
    for _ in range(max(1, args.warmup)):
        b.run()
    lfrv.device_sync()
    walls, lats = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        res = b.run()
        lfrv.device_sync()
        walls.append(time.perf_counter() - t)
        lats += [r["done_ms"] for r in res]
    identical = all(np.array_equal(ctx.get_solution(), s) for ctx, s in zip(ctxs, sols_serial))
    iters = [r["iterations"] for r in res]
    b.close()
    for c in ctxs:
        c.close()
    wall = float(np.median(walls))
    vps = n / wall
    vps_serial = n / t_serial
    line = {
        "metric": "novel views per second, batch of 64 uneven solves (target 256²-1536², 4-25 views, "
                  "PCG to rel. residual 1e-5 / 100 FISTA steps)",
        "value": round(vps, 2),
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched serving: 64 seeded solves, target side uniform in [256, 1536], 4-25 views, "
                               "half gradient-IBR PCG (dsf 1) and half IBR_color (dsf 2) from the tilted-plane "
                               "generator; one batched PCG launch per iteration replayed as a hipGraph",
                   "solves": n, "cg_solves": sum(p["kind"] == "cg" for p in problems),
                   "pd_solves": sum(p["kind"] == "pd" for p in problems), "seed": args.seed,
                   "autotuned": cfg},
        "batched": {"views_per_s": round(vps, 2), "wall_ms_median": round(1e3 * wall, 3),
                    "latency_p50_ms": round(float(np.percentile(lats, 50)), 3),
                    "latency_p99_ms": round(float(np.percentile(lats, 99)), 3)},
        "serial": {"views_per_s": round(vps_serial, 2), "wall_ms": round(1e3 * t_serial, 3),
                   "latency_p50_ms": round(float(np.percentile(lat_serial, 50)), 3),
                   "latency_p99_ms": round(float(np.percentile(lat_serial, 99)), 3)},
        "speedup_vs_serial": round(vps / vps_serial, 3),
        "solutions_identical_to_serial": bool(identical),
        "cg_iterations": {"min": int(min(i for i, p in zip(iters, problems) if p["kind"] == "cg")),
                          "max": int(max(i for i, p in zip(iters, problems) if p["kind"] == "cg"))},
        "autotune": {"seconds": round(t_tune, 1), "trials": len(trials),
                     "calibration": "the first 8 solves, 2 timed runs per configuration"},
    }
    print(json.dumps(line))


# ---- the learned warm start ----------------------------------------------------------

def conv_layer_rates(side, reps=20):
    """MFMA rate of the convolution kernel (lfrv_op_conv3x3, FP32 output) on
    the U-Net's layer shapes at a side x side input: 20 back-to-back launches
    per shape on random bf16 data, timed by the host around a device sync."""
    from lfrv.cnn import Raw, cnn_lib
    L = cnn_lib()
    rng = np.random.default_rng(5)
    shapes = [("enc0a", 0, 128, 32), ("enc0b", 0, 32, 32), ("enc1a", 1, 32, 64), ("enc1b", 1, 64, 64),
              ("enc2a", 2, 64, 128), ("enc2b", 2, 128, 128), ("enc3a", 3, 128, 256), ("enc3b", 3, 256, 256),
              ("dec2a", 2, 256, 128), ("dec1a", 1, 128, 64), ("dec0a", 0, 64, 32)]
    out = {}
    for name, lvl, ci, co in shapes:
        h = side >> lvl
        X = Raw((rng.integers(0x3f00, 0x3f80, h * h * ci, dtype=np.uint16)).astype(np.uint16))
        Wt = Raw((rng.integers(0x3c00, 0x3c80, co * 9 * ci, dtype=np.uint16)).astype(np.uint16))
        Y = Raw(h * h * co * 4)
        L.lfrv_op_conv3x3(1, h, h, ci, co, X.ptr, Wt.ptr, None, Y.ptr, None)
        lfrv.device_sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            L.lfrv_op_conv3x3(1, h, h, ci, co, X.ptr, Wt.ptr, None, Y.ptr, None)
        lfrv.device_sync()
        us = 1e6 * (time.perf_counter() - t0) / reps
        fl = 2.0 * 9 * ci * co * h * h
        out[name] = {"input": [h, h, ci], "cout": co, "us": round(us, 1), "tflops": round(fl / us * 1e-6, 1),
                     "frac": round(fl / us * 1e-6 / 2500.0, 4)}
        for b in (X, Wt, Y):
            b.free()
    return out


def bench_warm_start(args):
    """The learned warm start: 2000 scenes of the 5 x 5-view generator at its
    4x upsampling (64^2 views, 256^2 targets) solved to rel. residual 1e-5 by
    the plain PCG, the U-Net trained on them on the GPU, 200 held-out scenes
    solved both ways, then end-to-end timing at the full size (256^2 views,
    100 input channels at 1024^2).  One step = one full-size novel view,
    warm-started."""
    from lfrv import warmstart as wsm
    from lfrv.cnn import Cnn
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    n_train, n_test = args.train_scenes, args.test_scenes
    w_small, w_full, dsf = 64, 256, wsm.DSF
    t0 = time.perf_counter()
    data = wsm.TrainingSet(list(range(100000, 100000 + n_train)), w_small, log=log)
    t_data = time.perf_counter() - t0
    log(f"training set: {n_train} scenes in {t_data:.1f} s, plain PCG iterations mean {data.iters.mean():.1f}")
    net = Cnn(data.in_channels, levels=4, base=32, max_batch=8, max_h=data.H, max_w=data.W, seed=7)
    t0 = time.perf_counter()
    losses = wsm.train(net, data, epochs=args.epochs, batch=8, lr=1e-3, log=log)
    lfrv.device_sync()
    t_train = time.perf_counter() - t0
    data.free()
    params = net.get_params()
    res = wsm.compare(net, list(range(900000, 900000 + n_test)), w_small)
    net.close()
    ip = np.array([r["iters_plain"] for r in res], np.float64)
    iw = np.array([r["iters_warm"] for r in res], np.float64)
    used = np.array([r["used_net"] for r in res])
    rel = max(r["rel_diff"] for r in res)
    # full size: the trained weights on a 1024^2 network
    big = Cnn(data.in_channels, levels=4, base=32, max_batch=1, max_h=dsf * w_full, max_w=dsf * w_full, seed=7)
    big.set_params(params)
    seeds_full = list(range(990000, 990000 + args.steps + args.warmup))
    full = wsm.compare(big, seeds_full, w_full)[args.warmup:]
    # MFMA rate of the convolutions on one full-size forward
    big.set_timing(True)
    import ctypes
    from lfrv.cnn import Raw
    xin = Raw(dsf * dsf * w_full * w_full * big.cpad * 2)
    base = Raw(3 * dsf * dsf * w_full * w_full * 4)
    big.set_batch(1, dsf * w_full, dsf * w_full, [xin.ptr], [base.ptr])
    for _ in range(3):
        big.forward()
    conv_ms, conv_n, conv_flops = big.get_timing()
    big.set_timing(False)
    fwd_flops = big.forward_flops()
    xin.free()
    base.free()
    big.close()
    del ctypes
    tflops = conv_flops / (conv_ms * 1e-3) / 1e12
    layers = conv_layer_rates(dsf * w_full)
    ms_plain = float(np.median([r["ms_plain"] for r in full]))
    ms_warm = float(np.median([r["ms_warm"] for r in full]))
    line = {
        "metric": "ms per novel view, learned warm start + PCG to rel. residual 1e-5 (25 views 256² -> 1024²)",
        "value": round(ms_warm, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_warm, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network (FP32 accumulation), f32 solver",
        "data": "synthetic",
        "config": {"workload": "learned warm start: 4-level U-Net (32-256 channels, 3x3 convs, bf16 MFMA) on 25 "
                               "warped views + validity (100 channels) predicting the PCG start; trained on "
                               f"{n_train} solved scenes (64² views, 256² targets, 4x upsampling), tested on "
                               f"{n_test} held-out scenes and end to end at 256² views -> 1024²",
                   "train_scenes": n_train, "test_scenes": n_test, "epochs": args.epochs, "batch": 8},
        "held_out": {"scenes": n_test, "cg_iterations_plain_mean": round(float(ip.mean()), 3),
                     "cg_iterations_warm_mean": round(float(iw.mean()), 3),
                     "iterations_saved_mean": round(float((ip - iw).mean()), 3),
                     "iterations_saved_fraction": round(float(1 - iw.sum() / ip.sum()), 4),
                     "net_start_used": int(used.sum()), "max_rel_diff_vs_plain_solution": float("%.3g" % rel)},
        "full_size": {"views": len(full), "ms_plain_median": round(ms_plain, 3), "ms_warm_median": round(ms_warm, 3),
                      "speedup_vs_plain": round(ms_plain / ms_warm, 3),
                      "cg_iterations_plain": [r["iters_plain"] for r in full],
                      "cg_iterations_warm": [r["iters_warm"] for r in full],
                      "net_start_used": [r["used_net"] for r in full]},
        "training": {"seconds_data": round(t_data, 1), "seconds_train": round(t_train, 1),
                     "loss_first_epoch": float("%.4g" % losses[0]), "loss_last_epoch": float("%.4g" % losses[-1])},
        "roofline": {"bound": "mfma", "kernel": "k_conv3x3_halo / k_conv3x3 (implicit-GEMM 3x3 conv, bf16 MFMA 16x16x32)",
                     "achieved": round(tflops, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 4), "traffic": None,
                     "flops_per_forward": fwd_flops, "conv_launches": conv_n},
        "conv_layers": layers,
    }
    print(json.dumps(line))


def c2_step_bytes(N, Vn, e, Nc, iters):
    """Algorithmic HBM bytes of every kernel of one dsf = 1 gradient-IBR view
    (warps from depth to the PCG solution), each as its data must move once:
    inputs read, outputs written, per unit counts (N target pixels, Vn source
    pixels of all views, e binned entries, Nc footprint cells)."""
    return {
        "k_fill_key + k_key_to_depth": 8 * N + 12 * N,          # z-buffer keys set, read, depth written
        "k_tau_from_depth": 12 * Vn,                             # depth read; tau written (dtau/dz formed when read)
        "k_tau_visibility": 8 * Vn + 4 * N,                      # tau read, target depth read; counts
        "scan (2 kernels)": 12 * Nc,                             # counts read, starts and cursors written
        "k_cell_fill": 8 * Vn + 4 * e + 4 * Nc,                  # tau read, entry indices written, cursors
        "k_cell_sort": 4 * e + 12 * e + 16 * Nc,                 # indices read; sorted indices and tau written; tap sums
        "k_coverage_from_taps": 16 * Nc + 4 * N,
        "k_view_sobel": 20 * Vn,                                 # RGB read, gradients written
        "k_visibility": 8 * N,
        "k_weights<2,true>": 12 * Vn,                            # tau read, omega written
        "k_gibr_setup_st2": 24 * e + 84 * N,                     # entries (index, tau, omega) + view RGB / Sobel; 21 planes out
        "k_rhs_diag_cached": 32 * N,
        "k_cgcg_begin_st2 + k_cgcg_begin": 100 * N,
        "k_cgcg_st2": 144 * N * iters,                           # per PCG iteration
        "k_st2_residual": 48 * N,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile", type=str, default="", help="tile_w x tile_h of the binned splat kernels")
    ap.add_argument("--lds-kb", type=int, default=0, help="LDS budget of the pull kernels' entry cache (KiB)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the timed region, run the library's plain copy kernel once over MIB MiB "
                         "(reference dispatch for the HBM byte counters)")
    ap.add_argument("--workload", default="gradient_ibr_c2",
                    choices=["gradient_ibr_c2", "batch_c4", "warm_start_c5"] + sorted(COLOR_WORKLOADS),
                    help="gradient_ibr_c2 (the headline metric, default), batch_c4 (64 uneven solves at once) "
                         "or an IBR_color configuration")
    ap.add_argument("--no-tune-cache", action="store_true", help="batch_c4: autotune afresh, ignore the cache")
    ap.add_argument("--train-scenes", type=int, default=2000, help="warm_start_c5: training scenes")
    ap.add_argument("--test-scenes", type=int, default=200, help="warm_start_c5: held-out scenes")
    ap.add_argument("--epochs", type=int, default=16, help="warm_start_c5: training epochs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs one novel-view solve per GPU "
              "and every box has one GPU (multi-GPU sharding is a design note, not a measured path)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload in COLOR_WORKLOADS:
        bench_color_sr(args, args.workload)
        return
    if args.workload == "batch_c4":
        bench_batch(args)
        return
    if args.workload == "warm_start_c5":
        bench_warm_start(args)
        return

    scene = lfrv.synth(2, seed=args.seed)
    kw = {}
    if args.tile:
        tw, th = (int(v) for v in args.tile.lower().split("x"))
        kw = dict(tile_w=tw, tile_h=th)
    if args.lds_kb:
        kw["pull_lds_kb"] = args.lds_kb
    ctx = lfrv.Context(pipelines.gradient_ibr_params(scene, **kw))
    pipelines.load_scene(ctx, scene)
    ctx.sync()

    for _ in range(args.warmup):
        ctx.ibr_gradient(TOL, 5000)
    ctx.sync()

    ctx.set_timing(TIMING_STRIDE)
    iters = []
    per_step = []
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        ctx.set_target_camera(scene["P_target"])  # a novel view: its camera enters every step
        st = ctx.ibr_gradient(TOL, 5000)
        ctx.sync()
        per_step.append(1e3 * (time.perf_counter() - ts))
        iters.append(st["iterations"])
    ctx.sync()
    t1 = time.perf_counter()
    ms_per_view = 1e3 * (t1 - t0) / args.steps

    op_ms, op_n = ctx.get_timing("cg_op")
    up_ms, up_n = ctx.get_timing("cg_update")
    ctx.set_timing(False)
    U = ctx.get_solution()
    e_gpu = ctx.gibr_energy()
    st_last = st
    N = scene["W"] * scene["H"]
    tx = ctx.get_buffer("TAU_X")
    ty = ctx.get_buffer("TAU_Y")
    shp = (scene["V"], scene["h"], scene["w"])
    taus = tuple(ctx.get_buffer(b).reshape(shp) for b in ("TAU_X", "TAU_Y", "DPART_X", "DPART_Y"))
    e_unique = cell_valid_count(tx, ty, scene["W"], scene["H"])
    e_binned, ntiles = ctx.bin_stats()
    if scene["dsf"] == 1:
        # one fused single-reduction PCG iteration on the assembled 9-point
        # stencil, per target pixel: r, w, s_prev, p_prev, x (60 B), Minv (4 B)
        # and the 5 coefficient planes (20 B) read; p, x, s, r, w (60 B) written
        op_kernel = "k_cgcg_st2 (fused single-reduction PCG iteration, assembled 9-point stencil)"
        op_bytes = 144 * N
    else:
        # entry-walk operator: r, Minv, d_prev, d, q planes (52 B per target pixel)
        # + tau_x, tau_y, omega of every splatted source pixel (12 B)
        op_kernel = "k_cg_op (fused PCG operator, entry walk)"
        op_bytes = 52 * N + 12 * e_unique
    op_avg_s = op_ms / max(op_n, 1) / 1e3
    traffic = measured_traffic() if (scene["dsf"] == 1 and scene["V"] == 9 and N == 1024 * 1024) else None
    achieved = op_bytes / op_avg_s / 1e9
    up_bytes = 76 * N
    up_avg_s = up_ms / max(up_n, 1) / 1e3 if up_n else None
    psnr = lfrv.psnr(U, scene["target"])
    tile = [ctx.params.tile_w or 32, ctx.params.tile_h or 8]  # 0 = library default 32x8
    producers = warp_producer_times(ctx, scene)
    ctx.close()
    holes = hole_filling_times(scene)
    calib_copy(args)

    line = {
        "metric": METRIC,
        "value": round(ms_per_view, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_view, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "gradient-constraint IBR + Jacobi-PCG to rel. residual 1e-5: 9 views 1024x1024 "
                               "RGB, 8 tilted planes at depth 2-10, 3x3 cameras baseline 0.05 f=1000",
                   "views": scene["V"], "width": scene["W"], "height": scene["H"], "tol": TOL,
                   "seed": args.seed, "tile": tile},
        "roofline": {"bound": "hbm", "kernel": op_kernel,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": op_bytes, "avg_launch_us": round(op_avg_s * 1e6, 2),
                     "timed_launches": op_n, "timing_stride": TIMING_STRIDE},
        "ms_per_step_median": round(float(np.median(per_step)), 3),
        "ms_per_step_p10": round(float(np.percentile(per_step, 10)), 3),
        "ms_per_step_p90": round(float(np.percentile(per_step, 90)), 3),
        "cg_iterations": list(st_last["iterations"]),
        "true_rel_residual": [float("%.3g" % r) for r in st_last["true_rel_residual"]],
        "psnr_vs_ground_truth_db": round(psnr, 3),
        "kernels": {"binned_entries": e_binned, "unique_entries": e_unique, "tiles": ntiles},
        "warp_producers": producers,
        "hole_filling": holes,
    }
    if scene["dsf"] == 1:
        # the whole view against the HBM roofline: the algorithmic bytes of all
        # its kernels over the measured time per view
        sb = c2_step_bytes(N, scene["V"] * scene["w"] * scene["h"], e_binned, (scene["W"] + 1) * (scene["H"] + 1),
                           max(st_last["iterations"]))
        tot = sum(sb.values())
        gbs = tot / (ms_per_view * 1e-3) / 1e9
        line["step_roofline"] = {"bound": "hbm", "algorithmic_bytes": tot, "ms": round(ms_per_view, 3),
                                 "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(gbs / HBM_PEAK_GBS, 4),
                                 "bytes_by_kernel": {k: int(v) for k, v in sb.items()}}
    if up_n:
        line["kernels"]["cg_update"] = {"avg_launch_us": round(up_avg_s * 1e6, 2),
                                        "achieved_gbs": round(up_bytes / up_avg_s / 1e9, 1), "launches": up_n}
    if not args.no_cpu_baseline:
        threads = host_threads()
        line.update(error_vs_oracle(scene, dict(U=U, energy=e_gpu, psnr=psnr, taus=taus), threads))
        line["cpu_baseline"] = cpu_baseline(scene, threads)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
