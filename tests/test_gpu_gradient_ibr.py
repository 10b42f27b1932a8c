"""GPU parity of the gradient-domain IBR path (IBR_gradient) against the FP64 oracle."""
import numpy as np
import pytest

import lfrv
import oracle
from lfrv import pipelines
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _ctx(scene, **kw):
    ctx = lfrv.Context(pipelines.gradient_ibr_params(scene, **kw))
    pipelines.load_scene(ctx, scene)
    return ctx


def test_warps_from_depth_matches_oracle(small_scene):
    s = small_scene
    with _ctx(s) as ctx:
        ctx.warps_from_depth()
        tx = ctx.get_buffer("TAU_X").reshape(s["V"], s["h"], s["w"])
        ty = ctx.get_buffer("TAU_Y").reshape(s["V"], s["h"], s["w"])
        dpx = ctx.get_buffer("DPART_X").reshape(s["V"], s["h"], s["w"])
        zb = ctx.get_buffer("TARGET_DEPTH").reshape(s["H"], s["W"])
    otx, oty, odx, ody, oz = oracle.warps_from_depth(s["depths"], s["P_views"], s["P_target"], s["dsf"])
    valid = (otx >= 0) & (tx >= 0)
    # visibility decisions (int truncation of tau) may flip on a few boundary pixels
    assert np.mean((otx >= 0) != (tx >= 0)) < 2e-3
    assert np.max(np.abs(tx[valid] - otx[valid])) < 2e-3
    assert np.max(np.abs(ty[valid] - oty[valid])) < 2e-3
    assert np.max(np.abs(dpx - odx)) < 1e-3 * max(1.0, np.abs(odx).max())
    fin = oz < 1000
    assert np.mean(np.abs(zb[fin] - oz[fin]) > 1e-3) < 2e-3


def test_depth_partials_formed_when_read(small_scene):
    """IBR_gradient's import warps leave sigma dtau/dz to be formed when read:
    the planes read back are those of the warps' own depth even after a view's
    depth changes (formed before the change), equal to an IBR_color context's,
    which writes them with the warps; the weights that read them (a non-zero
    solution) match the eager context's too."""
    s = small_scene
    V, h, w = s["V"], s["h"], s["w"]
    print("depth partials: lazy (IBR_gradient) against eager (IBR_color)", flush=True)
    with _ctx(s) as ctx:
        ctx.warps_from_depth()
        ctx.set_depth(0, s["depths"][0] * 1.5, s["P_views"][0])
        dpx = ctx.get_buffer("DPART_X").reshape(V, h, w)
        dpy = ctx.get_buffer("DPART_Y").reshape(V, h, w)
    prm = pipelines.color_sr_params(s)
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.warps_from_depth()
        ex = ctx.get_buffer("DPART_X").reshape(V, h, w)
        ey = ctx.get_buffer("DPART_Y").reshape(V, h, w)
    assert np.array_equal(dpx, ex) and np.array_equal(dpy, ey)
    assert np.abs(ex).max() > 0.0
    # the weights of a non-zero solution read the planes: formed on that call
    U = np.clip(s["target"] + 0.05, 0.0, 1.0).astype(np.float32)
    om = []
    for lazy in (True, False):
        with _ctx(s) as ctx:
            ctx.warps_from_depth()
            if not lazy:
                ctx.get_buffer("DPART_X")  # formed before the weights ask
            ctx.prepare()
            ctx.set_solution(U)
            ctx.compute_weights()
            om.append(ctx.get_buffer("OMEGA"))
    assert np.array_equal(om[0], om[1])


@pytest.fixture(scope="module", params=["dsf1", "dsf2", "dsf3", "dsf4"])
def stage_scene(request, small_scene):
    if request.param == "dsf1":
        return small_scene      # 2x2 footprints: the CG operator runs as an assembled 9-point stencil
    # (dsf + 1)^2 footprints: the CG operator runs on the assembled anchor grid (box sums, 9-point couplings)
    if request.param == "dsf2":
        return lfrv.synth(2, seed=5, w=48, h=40, dsf=2)
    if request.param == "dsf3":
        return lfrv.synth(2, seed=6, w=40, h=32, dsf=3)
    return lfrv.synth(2, seed=7, w=32, h=24, dsf=4)


def test_pipeline_stages_match_oracle(stage_scene):
    """Every stage fed with the GPU's own warps, compared with the oracle stage."""
    s = stage_scene
    sig = 1.0 / 255.0
    with _ctx(s) as ctx:
        ctx.warps_from_depth()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        V, h, w, H, W = s["V"], s["h"], s["w"], s["H"], s["W"]
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        dpx = ctx.get_buffer("DPART_X").reshape(V, h, w)
        dpy = ctx.get_buffer("DPART_Y").reshape(V, h, w)
        om = ctx.get_buffer("OMEGA").reshape(V, h, w)
        vis = ctx.get_buffer("VISIBILITY").reshape(H, W)
        cov = ctx.get_buffer("COVERAGE").reshape(H, W)
        o_om = oracle.gibr_weights(W, H, s["dsf"], tx, ty, dpx, dpy, np.zeros((3, H, W)), sig, 1.0)
        assert rel_err(om, o_om) < 1e-6
        o_vis = oracle.target_visibility(W, H, s["dsf"], list(zip(tx, ty)))
        assert np.array_equal(vis.astype(np.uint8), o_vis)
        _, o_cov = oracle.splat(W, H, s["dsf"], tx[0], ty[0], None)
        for v in range(1, V):
            oracle.splat(W, H, s["dsf"], tx[v], ty[v], None, wsum=o_cov)
        assert np.max(np.abs(cov - o_cov)) < 1e-5
        ctx.compute_initial_image()
        U0 = ctx.get_solution()
        oU, _ = oracle.initial_image(W, H, s["dsf"], tx, ty, s["views"], om)
        oU, _ = oracle.hole_filling(oU, o_vis)
        assert np.max(np.abs(U0 - oU)) < 1e-5
        ctx.init_u_gradient()
        gx = ctx.get_buffer("U_GRAD_X").reshape(H, W)
        gy = ctx.get_buffer("U_GRAD_Y").reshape(H, W)
        tm = ctx.get_buffer("TARGET_MASK").reshape(H, W)
        ogx, ogy, otm, _ = oracle.u_gradient(W, H, s["dsf"], tx, ty, s["views"], om, sig)
        # dsf >= 2 splats in box form: FP32 sums of the same terms in another
        # order (2 x 2 anchor taps, then dsf x dsf anchors per pixel), so the
        # rounding bound grows with the dsf^2 anchors a pixel adds, and scales
        # with the terms (whose signs cancel in the gradient planes), not with
        # the result
        grow = max(1.0, s["dsf"] ** 2 / 4.0)
        assert np.max(np.abs(gx - ogx)) < 1e-6 * grow * max(1.0, np.abs(ogx).max())
        assert np.max(np.abs(gy - ogy)) < 1e-6 * grow * max(1.0, np.abs(ogy).max())
        assert np.max(np.abs(tm - otm)) < 4e-6 * grow * max(1.0, np.abs(otm).max())
        st = ctx.cg_solve(1e-5, 2000)
        U = ctx.get_solution()
        b = ctx.get_buffer("CG_B").reshape(3, H, W)
        wd = om.astype(np.float64) * sig * sig
        ob = oracle.rhs(W, H, s["dsf"], tx, ty, s["views"], wd, 1.0, 0.0, None, 0.5, gx, gy)
        assert rel_err(b, ob) < 1e-5
        oX, ost = oracle.pcg(W, H, s["dsf"], tx, ty, wd, 1.0, 0.0, 0.5, ob, U0, 1e-5, 2000)
        assert all(st["converged"]) and all(ost["converged"])
        assert max(st["true_rel_residual"]) < 2e-5
        assert rel_err(U, oX) < 1e-4
        E = ctx.gibr_energy()
        oE = oracle.gibr_energy(W, H, s["dsf"], tx, ty, s["views"], wd, 1.0, 0.0, None, 0.5, gx, gy, oX)
        assert abs(E - oE) / oE < 1e-4
        assert abs(oracle.psnr(U, s["target"]) - oracle.psnr(oX, s["target"])) < 0.01


def test_end_to_end_matches_oracle(small_scene):
    s = small_scene
    with _ctx(s) as ctx:
        st = pipelines.gradient_ibr_solve(ctx, 1e-5, 2000)
        U = ctx.get_solution()
    r = oracle.gibr_pipeline(s["views"], s["depths"], s["P_views"], s["P_target"], s["dsf"], 1 / 255.0, 1.0,
                             1.0, 0.5, 1e-5, 2000)
    assert all(st["converged"])
    assert rel_err(U, r["U"]) < 1e-3
    assert abs(oracle.psnr(U, s["target"]) - oracle.psnr(r["U"], s["target"])) < 0.01


def test_stateless_gather_splat_adjoint(small_scene):
    """<S u, r> == <u, S^T r> on the GPU's stateless operators (FP64 check of FP32 kernels)."""
    s = small_scene
    W, H, w, h, dsf = s["W"], s["H"], s["w"], s["h"], s["dsf"]
    rng = np.random.default_rng(0)
    tx = (np.arange(w)[None, :] + 0.5 + rng.uniform(0.5, 3.0, (h, w))).astype(np.float32)
    ty = (np.arange(h)[:, None] + 0.5 + rng.uniform(-2.0, 1.0, (h, w))).astype(np.float32)
    u = rng.standard_normal((3, H, W)).astype(np.float32)
    r = rng.standard_normal((3, h, w)).astype(np.float32)
    # restrict the gather to cell-valid pixels, like the splat
    cv = np.vectorize(lambda a, b: oracle.cell_valid(a, b, W, H))(tx, ty)
    txg = np.where(cv, tx, -1.0).astype(np.float32)
    D = lfrv.DeviceArray
    dtx, dty, dtxg, du, dr = D(tx), D(ty), D(txg), D(u), D(r)
    su, st = D((3, h, w)), D((3, H, W))
    L = lfrv.lib()
    assert L.lfrv_op_gather(W, H, w, h, dsf, dtxg.ptr, dty.ptr, du.ptr, 3, su.ptr, None) == 0
    assert L.lfrv_op_splat_atomic(W, H, w, h, dsf, dtx.ptr, dty.ptr, dr.ptr, 3, None, st.ptr, None, None) == 0
    lfrv.device_sync()
    su_h, st_h = su.numpy().astype(np.float64), st.numpy().astype(np.float64)
    lhs = float((su_h * r.astype(np.float64)).sum())
    rhs = float((u.astype(np.float64) * st_h).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    osu = oracle.gather(W, H, dsf, txg, ty, u)
    assert np.max(np.abs(su_h - osu)) < 1e-5
    ost, _ = oracle.splat(W, H, dsf, tx, ty, r)
    assert np.max(np.abs(st_h - ost)) < 1e-4


def test_sobel_gate_matches_oracle(small_scene):
    s = small_scene
    W, H = s["W"], s["H"]
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 1, (3, H, W)).astype(np.float32)
    vis = (rng.uniform(0, 1, (H, W)) > 0.05).astype(np.uint8)
    prev = rng.standard_normal((2, H, W)).astype(np.float32)
    D = lfrv.DeviceArray
    dimg, dgx, dgy = D(img), D(prev[0]), D(prev[1])
    dvis = lfrv.DeviceArray((H * W // 4 + 1,))
    vbytes = np.zeros(dvis.nbytes, np.uint8)
    vbytes[: H * W] = vis.ravel()
    assert lfrv.lib().lfrv_memcpy(dvis.ptr, vbytes.ctypes.data, dvis.nbytes) == 0
    assert lfrv.lib().lfrv_op_sobel(W, H, dimg.ptr, dgx.ptr, dgy.ptr, dvis.ptr, None) == 0
    lfrv.device_sync()
    ogx, ogy = oracle.sobel(img, vis, prev[0], prev[1])
    assert np.max(np.abs(dgx.numpy() - ogx)) < 1e-6
    assert np.max(np.abs(dgy.numpy() - ogy)) < 1e-6


@pytest.mark.parametrize("precond", [False, True])
def test_poisson_cg_matches_oracle(small_scene, precond):
    """The screened Poisson system of possion_conj_grad (solver.cu:361-505):
    (lambda I + D^T D) u = lambda u0 + D^T g from the hole-filled initial image."""
    s = small_scene
    lam = 1.0
    with _ctx(s) as ctx:
        ctx.warps_from_depth()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        ctx.init_u_gradient()
        H, W = s["H"], s["W"]
        U0 = ctx.get_solution().astype(np.float64)
        gx = ctx.get_buffer("U_GRAD_X").reshape(H, W)
        gy = ctx.get_buffer("U_GRAD_Y").reshape(H, W)
        tx = ctx.get_buffer("TAU_X").reshape(s["V"], s["h"], s["w"])
        ty = ctx.get_buffer("TAU_Y").reshape(s["V"], s["h"], s["w"])
        st = ctx.poisson_cg(lam, 1e-6, 3000, precond=precond)
        U = ctx.get_solution()
    wz = np.zeros(tx.shape)
    ob = oracle.rhs(W, H, s["dsf"], tx, ty, s["views"], wz, 0.0, lam, U0, 1.0, gx, gy)
    oX, ost = oracle.pcg(W, H, s["dsf"], tx, ty, wz, 0.0, lam, 1.0, ob, U0, 1e-6, 3000, precond)
    assert all(st["converged"]) and all(ost["converged"])
    assert max(st["true_rel_residual"]) < 5e-6
    assert rel_err(U, oX) < 1e-5


def _holed_scene(scene):
    """The small scene with a block of invalid depth (0 < 1e-3) in every view:
    the target pixels only those source pixels reached become holes."""
    s = dict(scene)
    d = scene["depths"].copy()
    d[:, 18:38, 20:44] = 0.0
    s["depths"] = d
    return s


def test_hole_filling_matches_oracle(small_scene):
    s = _holed_scene(small_scene)
    V, h, w, H, W = s["V"], s["h"], s["w"], s["H"], s["W"]
    with _ctx(s) as ctx:
        ctx.clear_solution()
        ctx.warps_from_depth()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        U0 = ctx.get_solution()
        nholes, sweeps = ctx.hole_stats()
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        om = ctx.get_buffer("OMEGA").reshape(V, h, w)
        vis = ctx.get_buffer("VISIBILITY").reshape(H, W).astype(np.uint8)
        # the same image again with the hole list rebuilt from a visibility set by hand
        ctx.set_buffer("VISIBILITY", vis.astype(np.float32))
        ctx.compute_initial_image()
        U1 = ctx.get_solution()
    assert (vis == 0).sum() > 100           # the scene has holes
    oU, _ = oracle.initial_image(W, H, s["dsf"], tx, ty, s["views"], om)
    oU, k = oracle.hole_filling(oU, vis)
    assert k > 1
    assert nholes == int((vis == 0).sum()) and abs(sweeps - k) <= 1
    # Jacobi sweeps to an energy-change test: FP32 and FP64 may stop one sweep apart
    assert np.max(np.abs(U0 - oU)) < 2e-3
    assert np.max(np.abs(U0[:, vis == 1] - oU[:, vis == 1])) < 1e-5
    assert np.array_equal(U0, U1)


@pytest.mark.parametrize("holes", [False, True])
def test_ibr_gradient_call_equals_the_sequence(small_scene, holes):
    """lfrv_ibr_gradient gives the results of the calls it stands for, made one by one."""
    s = _holed_scene(small_scene) if holes else small_scene
    with _ctx(s) as ctx:
        st_seq = pipelines.gradient_ibr_solve(ctx, 1e-5, 2000)
        U_seq = ctx.get_solution()
    with _ctx(s) as ctx:
        st_one = ctx.ibr_gradient(1e-5, 2000)
        U_one = ctx.get_solution()
        st_two = ctx.ibr_gradient(1e-5, 2000)  # a second solve on the same context (cached set-up)
        U_two = ctx.get_solution()
    assert st_one["iterations"] == st_seq["iterations"] == st_two["iterations"]
    assert all(st_one["converged"])
    assert np.max(np.abs(U_one - U_seq)) <= 1e-6
    assert np.array_equal(U_one, U_two)
