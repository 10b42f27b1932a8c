"""Failure handling on the GPU.  Every condition is forced from the host, so
every kernel still runs to its end:

* NaN in the input views (reported as E_NAN, not as a solution);
* an operator that is not positive definite: the CG breakdown that
  possion_conj_grad asserts on (a zero or negative d.q, solver.cu:630, 662) is
  reported as E_BREAKDOWN, for the fused stencil PCG (dsf 1) and for the
  entry-walk PCG (dsf 2); a negative screening weight is refused (E_ARG);
* an iteration cap below what the tolerance needs (E_NONCONVERGED, with the
  iterations it ran);
* calls out of order (E_STATE) and arguments out of range (E_ARG), with the
  context still usable afterwards;
* a scene whose depths are all invalid (import.cpp:292-296): no warp reaches
  the target, zero coverage everywhere, every pixel a hole, and the solve
  ends at once on a zero right-hand side;
* the init_name starting image of the command line: read when the PFM exists
  and has the target's size, else the splat average is computed
  (gradientIBR.cu:737-762)."""
import numpy as np
import pytest

import lfrv
from lfrv import io as lio
from lfrv import pipelines
from conftest import rel_err
from test_gpu_cli import _config, _run_cli, _write_scene

pytestmark = pytest.mark.gpu


def _ctx(scene, **kw):
    ctx = lfrv.Context(pipelines.gradient_ibr_params(scene, **kw))
    pipelines.load_scene(ctx, scene)
    return ctx


def test_nan_in_the_views_is_reported(small_scene):
    s = dict(small_scene)
    views = s["views"].copy()
    c = s["V"] // 2  # the centre view lands next to the target
    views[c][:, 30:34, 30:34] = np.nan
    s["views"] = views
    print("NaN block in the centre view", flush=True)
    with _ctx(s) as ctx:
        st = ctx.ibr_gradient(1e-5, 200, check=False)
        assert st["status"] == lfrv.E_NAN
        assert not any(st["converged"])
        with pytest.raises(lfrv.LfrvError) as ei:
            ctx.ibr_gradient(1e-5, 200)
        assert ei.value.status == lfrv.E_NAN
        # the same context solves the clean views afterwards
        ctx.set_view(c, small_scene["views"][c])
        st = ctx.ibr_gradient(1e-5, 2000)
        assert st["status"] == lfrv.OK and all(st["converged"])
        assert np.all(np.isfinite(ctx.get_solution()))


def test_breakdown_of_the_stencil_pcg(small_scene):
    """dt_alpha < 0 and no gradient term: A = dt_alpha S^T W S is negative
    semi-definite, so the first u.Au is not positive."""
    print("IBR_gradient with dt_alpha = -1, dt_beta = 0", flush=True)
    with _ctx(small_scene, dt_alpha=-1.0, dt_beta=0.0) as ctx:
        st = ctx.ibr_gradient(1e-5, 100, check=False)
        assert st["status"] == lfrv.E_BREAKDOWN
        assert st["iterations"] == [0, 0, 0]
        assert np.all(np.isfinite(ctx.get_solution()))
        with pytest.raises(lfrv.LfrvError) as ei:
            ctx.ibr_gradient(1e-5, 100)
        assert ei.value.status == lfrv.E_BREAKDOWN


def test_breakdown_of_the_anchor_grid_pcg():
    """dsf 2 (3 x 3 footprints): the operator is applied on the assembled
    anchor grid (k_box_pass) and the first alpha taken by the begin step;
    dt_alpha < 0 without the gradient term makes w.u negative on the first
    direction."""
    s = lfrv.synth(2, seed=5, w=48, h=40, dsf=2)
    print("IBR_gradient dsf 2 with dt_alpha = -1, dt_beta = 0", flush=True)
    with _ctx(s, dt_alpha=-1.0, dt_beta=0.0) as ctx:
        ctx.clear_solution()
        ctx.warps_from_depth()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        ctx.init_u_gradient()
        U0 = ctx.get_solution()
        st = ctx.cg_solve(1e-5, 100, check=False)
        assert st["status"] == lfrv.E_BREAKDOWN
        assert st["iterations"] == [0, 0, 0]
        # alpha is never taken: the solution is the starting point
        assert np.array_equal(ctx.get_solution(), U0)


def test_poisson_cg_rejects_a_negative_lambda(small_scene):
    """lambda I + D^T D is positive definite only for lambda >= 0; a negative
    lambda is refused before any work, and the status says so also when the
    caller does not raise on it."""
    print("poisson_cg lambda = -10", flush=True)
    with _ctx(small_scene) as ctx:
        pipelines.gradient_ibr_solve(ctx)
        U0 = ctx.get_solution()
        st = ctx.poisson_cg(-10.0, 1e-6, 100, check=False)
        assert st["status"] == lfrv.E_ARG
        with pytest.raises(lfrv.LfrvError) as ei:
            ctx.poisson_cg(-10.0, 1e-6, 100)
        assert ei.value.status == lfrv.E_ARG
        assert np.array_equal(ctx.get_solution(), U0)


def test_iteration_cap_is_reported(small_scene):
    print("IBR_gradient capped at 2 iterations", flush=True)
    with _ctx(small_scene) as ctx:
        st = ctx.ibr_gradient(1e-12, 2, check=False)
        assert st["status"] == lfrv.E_NONCONVERGED
        assert st["iterations"] == [2, 2, 2] and not any(st["converged"])
        assert min(st["rel_residual"]) > 1e-12
        with pytest.raises(lfrv.LfrvError) as ei:
            ctx.ibr_gradient(1e-12, 2)
        assert ei.value.status == lfrv.E_NONCONVERGED
        # a cap the tolerance fits under converges in the same context
        st = ctx.ibr_gradient(1e-5, 2000)
        assert all(st["converged"]) and max(st["true_rel_residual"]) < 1e-5


@pytest.mark.parametrize("dsf", [1, 2])
def test_zero_iteration_cap(small_scene, dsf):
    """maxit = 0 runs no iteration: the starting point comes back unchanged
    with its own residual (recursive and true agree: r_0 = b - A x_0), the
    status says it did not converge, and the next solve in the context is
    the same as in a fresh one.  dsf 1: the stencil path; dsf 2: the anchor
    grid."""
    print(f"IBR_gradient with maxit 0 after a full solve, dsf {dsf}", flush=True)
    if dsf == 2:
        small_scene = lfrv.synth(2, seed=5, w=48, h=40, dsf=2)
    with _ctx(small_scene) as ctx:
        ctx.ibr_gradient(1e-5, 2000)           # the context has a previous solve behind it
        ctx.clear_solution()
        ctx.warps_from_depth()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        ctx.init_u_gradient()
        U0 = ctx.get_solution()
        st = ctx.cg_solve(1e-5, 0, check=False)
        assert st["status"] == lfrv.E_NONCONVERGED
        assert st["iterations"] == [0, 0, 0] and not any(st["converged"])
        assert np.array_equal(ctx.get_solution(), U0)
        for c in range(3):
            assert st["rel_residual"][c] > 1e-5
            assert abs(st["true_rel_residual"][c] - st["rel_residual"][c]) <= 1e-3 * st["rel_residual"][c]
        st = ctx.cg_solve(1e-5, 2000)
        U1 = ctx.get_solution()
    with _ctx(small_scene) as ctx:
        ctx.ibr_gradient(1e-5, 2000)
        assert np.array_equal(ctx.get_solution(), U1)
    assert all(st["converged"])


def test_calls_out_of_order_and_bad_arguments(small_scene):
    s = small_scene
    print("state and argument errors", flush=True)
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        with pytest.raises(lfrv.LfrvError) as ei:
            ctx.warps_from_depth()           # no depth, no target camera yet
        assert ei.value.status == lfrv.E_STATE
        with pytest.raises(lfrv.LfrvError) as ei:
            ctx.setup_visibility()           # before the splat structure exists
        assert ei.value.status == lfrv.E_STATE
        with pytest.raises(lfrv.LfrvError) as ei:
            ctx.set_view(s["V"], s["views"][0])
        assert ei.value.status == lfrv.E_ARG
        with pytest.raises(lfrv.LfrvError) as ei:
            ctx.set_view(-1, s["views"][0])
        assert ei.value.status == lfrv.E_ARG
        # the context is still good for the whole solve
        pipelines.load_scene(ctx, s)
        st = ctx.ibr_gradient(1e-5, 2000)
        assert all(st["converged"])


def test_all_depths_invalid(small_scene):
    s = dict(small_scene)
    s["depths"] = np.zeros_like(small_scene["depths"])   # d < 1e-3 everywhere
    print("every depth invalid", flush=True)
    with _ctx(s) as ctx:
        st = ctx.ibr_gradient(1e-5, 2000)
        tx = ctx.get_buffer("TAU_X")
        cov = ctx.get_buffer("COVERAGE")
        U = ctx.get_solution()
        holes, _ = ctx.hole_stats()
    assert np.all(tx < 0)
    assert np.all(cov == 0)
    assert st["status"] == lfrv.OK and all(st["converged"]) and st["iterations"] == [0, 0, 0]
    # every target pixel is a hole; diffusion from no known pixel leaves zeros
    assert holes == s["W"] * s["H"]
    assert np.all(U == 0)


def _png(U):
    return np.clip(np.trunc(U.astype(np.float64) * 255.0), 0, 255) / 255.0


def test_cli_init_name(tmp_path, small_scene):
    s = small_scene
    s_min, s_max = _write_scene(tmp_path, s)
    base = ["dt_alpha 1.0", "dt_beta 0.5", "ibr_system normal_equations"]
    runs = {}
    rng = np.random.default_rng(5)
    init = rng.uniform(0.0, 1.0, (3, s["H"], s["W"])).astype(np.float32)
    lio.save_image(str(tmp_path / "init.pfm"), init)
    lio.save_image(str(tmp_path / "small.pfm"), init[:, :-1, :])
    for name, extra in (("none", []), ("missing", [f"init_name {tmp_path}/absent.pfm"]),
                        ("wrong_size", [f"init_name {tmp_path}/small.pfm"]),
                        ("init", [f"init_name {tmp_path}/init.pfm"])):
        d = tmp_path / name
        d.mkdir()
        cfg = _config(tmp_path, s, s_min, s_max, -1, "IBR_gradient", base + extra)
        text = cfg.read_text().replace(f"outdir {tmp_path}/out", f"outdir {d}")
        cfg.write_text(text)
        _run_cli(cfg)
        runs[name] = (lio.load_image(str(d / "output.pfm")), lio.load_image(str(d / "u_init_filled.png")))
    # an unreadable or mis-sized file: the computed starting image, the same solution
    for name in ("missing", "wrong_size"):
        assert np.array_equal(runs[name][0], runs["none"][0])
        assert np.array_equal(runs[name][1], runs["none"][1])
    # a readable one is the starting image: the run equals the C-ABI sequence
    # started from it
    assert np.max(np.abs(runs["init"][1] - _png(init))) <= 1.0 / 255.0 + 1e-7
    assert not np.array_equal(runs["init"][1], runs["none"][1])
    prm = pipelines.gradient_ibr_params(s, dt_alpha=1.0, dt_beta=0.5, sigma_sensor=np.float32(1.0) / np.float32(255))
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.clear_solution()
        ctx.warps_from_depth()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image(init)
        assert np.array_equal(ctx.get_solution(), init)
        ctx.init_u_gradient()
        st = ctx.cg_solve(1e-5, 5000)
        U = ctx.get_solution()
    assert all(st["converged"])
    assert np.max(np.abs(runs["init"][0] - U)) <= 1e-6 * max(1.0, np.abs(U).max())
    assert rel_err(runs["init"][0], runs["none"][0]) < 1e-2
