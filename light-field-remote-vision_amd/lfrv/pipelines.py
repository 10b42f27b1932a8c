"""Host-side sequences of the two reference drivers over the C-ABI.

gradient_ibr follows IBR_gradient (sources/gradientIBR.cpp:12-150) with the
warps computed from depth on the GPU and the CG of the normal equations;
color_sr_setup follows the set-up half of sr_synthesize_view
(sources/super-resolution.cpp:54-166).
"""
import numpy as np

from . import MODE_COLOR_SR, MODE_GRADIENT_IBR, Context, default_params


def load_scene(ctx, scene, with_depth=True):
    for i in range(scene["V"]):
        ctx.set_view(i, scene["views"][i])
        if with_depth:
            ctx.set_depth(i, scene["depths"][i], scene["P_views"][i])
    if with_depth:
        ctx.set_target_camera(scene["P_target"])


def gradient_ibr_params(scene, **kw):
    p = dict(dt_alpha=1.0, dt_beta=0.5)
    p.update(kw)
    return default_params(MODE_GRADIENT_IBR, scene["V"], scene["w"], scene["h"], scene["dsf"], **p)


def gradient_ibr_solve(ctx, tol=1e-5, maxit=2000, check=True):
    """Device-resident part of IBR_gradient: from the views, depths and cameras
    already in the context to the solution in ctx.U."""
    ctx.clear_solution()
    ctx.warps_from_depth()
    ctx.prepare()
    ctx.setup_visibility()
    ctx.compute_weights()
    ctx.compute_initial_image()
    ctx.init_u_gradient()
    return ctx.cg_solve(tol, maxit, check=check)


def color_sr_params(scene, **kw):
    p = dict(lambda_=0.05, lambda_max_factor=100.0, inner_iterations=1, dt_alpha=1.0, dt_beta=0.0)
    p.update(kw)
    lam = p.pop("lambda_")
    return default_params(MODE_COLOR_SR, scene["V"], scene["w"], scene["h"], scene["dsf"], **{"lambda": lam}, **p)


def color_sr_setup(ctx):
    """sr_synthesize_view up to FINISHED INIT (super-resolution.cpp:106-166)."""
    ctx.warps_from_depth()
    ctx.prepare()
    ctx.compute_weights()
    ctx.setup_visibility()
    ctx.compute_initial_image()
