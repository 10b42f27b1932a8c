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


def _scene(w=48, h=40, seed=4, dsf=1):
    d = lfrv.synth_desc(2, seed)
    d.w, d.h = w, h
    d.focal = 1.5 * w      # disparities up to a few pixels, with occlusions between the planes
    d.dsf = d.render_scale = dsf
    return lfrv.synth(d)


def _grads(s, seed):
    r = np.random.default_rng(seed)
    return [(r.uniform(-0.02, 0.02, (s["h"], s["w"])).astype(np.float32),
             r.uniform(-0.02, 0.02, (s["h"], s["w"])).astype(np.float32)) for _ in range(s["V"])]


# dsf 1 draws one quad per thread (k_gl_splat); dsf >= 2 deals the quads'
# columns to the lanes of a wave (k_gl_splat_wave)
@pytest.mark.parametrize("with_grad,skip,dsf", [(False, -1, 1), (True, -1, 1), (False, 4, 1), (False, -1, 2),
                                                (True, 4, 3)])
def test_gl_producer_matches_oracle(with_grad, skip, dsf):
    s = _scene(dsf=dsf)
    V, h, w = s["V"], s["h"], s["w"]
    W, H = s["W"], s["H"]
    g = _grads(s, 9) if with_grad else None
    print(f"gl producer grad={with_grad} skip={skip} dsf={dsf}", flush=True)
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


def test_export_warps_files_feed_the_cli(tmp_path):
    """export_warps (Scene::computeTargetDepthMap + exportWarps) writes the tau /
    dpart / uDepth PFMs of the OpenGL path; the command-line IBR_gradient reads
    them back (read_tau, read_partial_tau) and solves as the in-process sequence
    on the same warps does."""
    import subprocess
    import sys
    from lfrv import io as lio
    s = _scene(40, 32, seed=8)
    V, h, w = s["V"], s["h"], s["w"]
    rmv = 4                                  # the target is the centre view of the 3 x 3 grid
    for v in range(V):
        lio.save_image(str(tmp_path / f"v_{v:02d}.pfm"), s["views"][v])
        lio.save_image(str(tmp_path / f"d_{v:02d}.pfm"), np.repeat(s["depths"][v][None], 3, 0))
        with open(tmp_path / f"cam_{v:02d}.txt", "w") as f:
            f.write("CONTOUR\n" + "\n".join(" ".join(f"{x:.17g}" for x in r) for r in s["P_views"][v]) + "\n")
    lio.export_warps(["export", "-depth_name", f"{tmp_path}/d_%02d.pfm", "-cam_name", f"{tmp_path}/cam_%02d.txt",
                      "-tau_name", f"{tmp_path}/t_%02d.pfm", "-dpart_name", f"{tmp_path}/p_%02d.pfm",
                      "-uDepth_name", f"{tmp_path}/ud.pfm", "-s_min", "0", "-s_max", str(V - 1), "-s_rmv", str(rmv),
                      "-dsf", "1", "-outdir", str(tmp_path)])
    # the files hold the producer's buffers
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        pipelines.load_scene(ctx, s, with_depth=False)
        for v in range(V):
            ctx.set_depth(v, s["depths"][v], s["P_views"][v])
        ctx.set_target_camera(s["P_views"][rmv])
        ctx.warps_from_target_depth(rmv)
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        wd = ctx.get_buffer("TAU_WEIGHT").reshape(V, h, w)
        dpy = ctx.get_buffer("DPART_Y").reshape(V, h, w)
        ud = ctx.get_buffer("TARGET_DEPTH").reshape(h, w)
    t3 = lio.load_image(str(tmp_path / "t_03.pfm"))
    assert np.array_equal(t3[0], tx[3]) and np.array_equal(t3[2], wd[3])
    p3 = lio.load_image(str(tmp_path / "p_03.pfm"))
    assert np.all(p3[0] == 10.0) and np.allclose(p3[2] * 10.0, dpy[3], rtol=1e-6, atol=1e-9)
    assert np.array_equal(lio.load_image(str(tmp_path / "ud.pfm"))[0], ud)
    # the target view's own warp is the identity where it sees itself
    X = np.arange(w) + 0.5
    ok = tx[rmv] > 0
    assert ok.mean() > 0.8 and np.max(np.abs(tx[rmv][ok] - np.broadcast_to(X, (h, w))[ok])) < 1e-3
    # the command line reads those files back (import_warps + read_tau / read_partial_tau) for every view
    # but the target and solves as the same calls made in this process
    cfg = tmp_path / "gibr.cfg"
    cfg.write_text("\n".join([
        "algorithm IBR_gradient", f"lf_name {tmp_path}/v_%02d.pfm", f"tau_name {tmp_path}/t_%02d.pfm",
        f"dpart_name {tmp_path}/p_%02d.pfm", f"outdir {tmp_path}/out", "outfile u.pfm", "s_min 0",
        f"s_max {V - 1}", f"s_rmv {rmv}", "dsf 1", "nchannels 3", "sigma_sensor 1", "dt_beta 0.5",
        "ibr_system normal_equations"]) + "\n")
    r = subprocess.run([lio.CLI_PATH, "-config", str(cfg)], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr
    U_cli = lio.load_image(str(tmp_path / "out" / "u.pfm"))
    keep = [v for v in range(V) if v != rmv]
    s2 = dict(s, V=len(keep), views=s["views"][keep])
    prm = pipelines.gradient_ibr_params(s2, dt_beta=0.5, sigma_sensor=np.float32(1.0) / np.float32(255))
    with lfrv.Context(prm) as ctx:
        for i, v in enumerate(keep):
            ctx.set_view(i, s["views"][v])
            t = lio.load_image(str(tmp_path / f"t_{v:02d}.pfm"))
            p = lio.load_image(str(tmp_path / f"p_{v:02d}.pfm"))
            ctx.set_tau(i, t[0], t[1])
            ctx.set_partial_tau(i, p[0], p[1], p[2])
        ctx.clear_solution()
        ctx.prepare()
        ctx.setup_visibility()
        ctx.compute_weights()
        ctx.compute_initial_image()
        ctx.init_u_gradient()
        st = ctx.cg_solve(1e-5, 5000)
        U = ctx.get_solution()
    assert all(st["converged"])
    assert np.max(np.abs(U_cli - U)) <= 1e-6 * max(1.0, np.abs(U).max())


@pytest.mark.parametrize("dsf,zoom", [(1, 90.0), (4, 30.0)])
def test_gl_producer_rasterises_big_quads(dsf, zoom):
    """A target camera magnifying `zoom` times beyond the grid's: each source
    pixel's quad spans more target pixels than the per-thread walk takes (64),
    so it goes through the overflow pass; the target depth still matches the
    oracle (which rasterises every quad)."""
    s = _scene(w=40, h=32, seed=6, dsf=dsf)
    W, H = s["W"], s["H"]
    Pu = s["P_target"].copy()
    Pu[:2, :] *= zoom                   # K_u scaled about the principal point: x_u - c = zoom (x - c)
    Pu[0, :] += (1 - zoom) * Pu[2, :] * (W / 2)
    Pu[1, :] += (1 - zoom) * Pu[2, :] * (H / 2)
    print(f"GL producer, big quads: dsf {dsf}, zoom {zoom}", flush=True)
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.set_target_camera(Pu)
        ctx.warps_from_target_depth(-1)
        ud = ctx.get_buffer("TARGET_DEPTH").reshape(H, W)
    oud = oracle.gl_target_depth(s["depths"], s["P_views"], Pu, W, H)
    assert (oud > 0).mean() > 0.5
    bad = np.abs(ud - oud) > 1e-4 * np.maximum(np.abs(oud), 1.0)
    assert bad.mean() <= 5e-3, bad.mean()


def test_radial_to_ortho_matches_oracle():
    """The depth-map filter of the OpenGL path (fromRadial2Ortho.frag) on the
    GPU against the oracle, on random depths and derivatives with a skewed K."""
    r = np.random.default_rng(12)
    H, W = 37, 53
    K = np.array([[70.0, 1.5, 25.0], [0.0, 64.0, 19.0], [0.0, 0.0, 1.0]])
    z = r.uniform(1.0, 9.0, (H, W)).astype(np.float32)
    z[3, 4] = 1000.0                                   # an invalid pixel is converted too
    gx = r.normal(size=(H, W)).astype(np.float32)
    gy = r.normal(size=(H, W)).astype(np.float32)
    zg, gxg, gyg = lfrv.radial_to_ortho(z, K, gx, gy)
    zo, gxo, gyo = oracle.radial_to_ortho(z, K, gx, gy)
    assert np.allclose(zg, zo, rtol=2e-6)
    assert np.allclose(gxg, gxo, rtol=1e-5, atol=1e-5) and np.allclose(gyg, gyo, rtol=1e-5, atol=1e-5)
    z2, n1, n2 = lfrv.radial_to_ortho(z, K)            # the derivatives are optional
    assert np.array_equal(z2, zg) and n1 is None and n2 is None


def test_from_mve_to_coco(tmp_path):
    """The fromMVE2coco algorithm through the command line: a synthetic scene
    written as an MVE scene (meta.ini cameras, radial depth maps, zero dz,
    undistorted.png), imported, filtered to orthographic depth on the GPU and
    turned into warps.  The filtered depth is the scene's own (orthographic)
    depth; the files hold what lfrv_warps_from_target_depth gives on the
    filtered maps, and those warps agree with the oracle's producer; a view
    without a camera is left out."""
    import subprocess
    import sys
    from lfrv import io as lio
    from mve_files import decompose, mve_params, radial_from_ortho, view_dir, write_meta, write_mvei
    s = _scene(40, 32, seed=8)
    V, h, w = s["V"], s["h"], s["w"]
    rmv = 4
    for v in range(V + 1):                              # view V has no camera
        d = view_dir(tmp_path, v)
        src = min(v, V - 1)
        lio.save_image(f"{d}/undistorted.png", s["views"][src])
        K, R, t, sc = decompose(s["P_views"][src])
        assert abs(sc - 1.0) < 1e-12 and abs(K[0, 1]) < 1e-12
        f, a, pp = mve_params(K, w, h)
        write_meta(f"{d}/meta.ini", f, a, pp, R, t, with_camera=v < V)
        zr = radial_from_ortho(s["depths"][src].astype(np.float64), K).astype(np.float32)
        write_mvei(f"{d}/depth-L0.mvei", zr[None])
        write_mvei(f"{d}/dz-L0.mvei", np.zeros((2, h, w), np.float32))
    cfg = tmp_path / "mve.cfg"
    cfg.write_text("\n".join([
        "algorithm fromMVE2coco", f"mve_name {tmp_path}", f"lf_name {tmp_path}/v_%02d.png",
        f"tau_name {tmp_path}/t_%02d.pfm", f"dpart_name {tmp_path}/p_%02d.pfm", f"uDepth_name {tmp_path}/ud.pfm",
        f"outdir {tmp_path}/out", f"width {w}", f"height {h}", "s_min 0", f"s_max {V}", f"s_rmv {rmv}", "dsf 1",
        "scale_min 0", "scale_max 0"]) + "\n")
    r = subprocess.run([lio.CLI_PATH, "-config", str(cfg)], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stderr
    assert not (tmp_path / f"t_{V:02d}.pfm").exists()   # V views imported, numbered 0..V-1
    # the same steps in process: the filtered depth is the scene's depth
    gxs, gys = [], []
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        for v in range(V):
            zr, gx, gy = lio.mve_depth(str(tmp_path), v, 0, 0, w, h)
            K, _, _, P = lio.mve_camera(f"{view_dir(tmp_path, v)}/meta.ini", w, h)
            zo, gxo, gyo = lfrv.radial_to_ortho(zr, K, gx, gy)
            assert np.allclose(zo, s["depths"][v], rtol=1e-5)
            ctx.set_depth(v, zo, P)
            ctx.set_depth_gradient(v, gxo, gyo)
            gxs.append(gxo)
            gys.append(gyo)
        _, _, _, Pt = lio.mve_camera(f"{view_dir(tmp_path, rmv)}/meta.ini", w, h)
        assert np.allclose(Pt, s["P_views"][rmv], rtol=1e-12, atol=1e-9)
        ctx.set_target_camera(Pt)
        ctx.warps_from_target_depth(rmv)
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        ud = ctx.get_buffer("TARGET_DEPTH").reshape(h, w)
    for v in (0, 3, 8):
        t = lio.load_image(str(tmp_path / f"t_{v:02d}.pfm"))
        assert np.array_equal(t[0], tx[v]) and np.array_equal(t[1], ty[v])
        assert np.array_equal(lio.load_image(str(tmp_path / f"v_{v:02d}.png")),
                              lio.load_image(f"{view_dir(tmp_path, v)}/undistorted.png"))
    assert np.array_equal(lio.load_image(str(tmp_path / "ud.pfm"))[0], ud)
    # those warps against the oracle's producer on the scene's own depth and the filtered derivatives
    gx, gy = np.stack(gxs), np.stack(gys)
    oud = oracle.gl_target_depth(s["depths"], s["P_views"], s["P_views"][rmv], w, h, rmv, gx, gy)
    assert (np.abs(ud - oud) > 1e-3 * np.maximum(1.0, oud)).mean() < 5e-3
    otx, oty, _, _, _ = oracle.gl_warps(s["depths"], s["P_views"], s["P_views"][rmv], ud, 10.0, gx, gy)
    both = (tx >= 0) & (otx >= 0)
    assert ((tx >= 0) == (otx >= 0)).mean() > 0.998
    assert np.max(np.abs(tx[both] - otx[both])) < 2e-3 and np.max(np.abs(ty[both] - oty[both])) < 2e-3
