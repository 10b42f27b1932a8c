"""The warm-start U-Net on the GPU against its FP64 oracle: single 3x3
convolutions and weight gradients on the bf16 MFMA path, the whole
network's prediction, loss and parameter gradients (the oracle rounding
what the GPU stores as bf16), training, the warped-view stack, and the warm
start's energy rule on a solver context."""
import ctypes

import numpy as np
import pytest

import lfrv
import oracle
from lfrv import pipelines
from lfrv.cnn import Cnn, Raw, WarmStats, cnn_lib, from_bf16, layer_slices, to_bf16
from conftest import rel_err

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(2026)


@pytest.mark.parametrize("B,H,W,Cin,Cout", [(1, 5, 7, 32, 3), (2, 9, 13, 64, 48), (1, 16, 16, 128, 32),
                                            (1, 8, 8, 32, 256), (3, 4, 4, 256, 64)])
def test_conv3x3_matches_oracle(B, H, W, Cin, Cout):
    print(f"conv3x3 B={B} {H}x{W} {Cin}->{Cout}", flush=True)
    L = cnn_lib()
    X = to_bf16(RNG.standard_normal((B, H, W, Cin)).astype(np.float32))
    Wt = to_bf16((RNG.standard_normal((Cout, 9, Cin)) / np.sqrt(9 * Cin)).astype(np.float32))
    b = RNG.standard_normal(Cout).astype(np.float32)
    dX, dW, db = Raw(X), Raw(Wt), Raw(b)
    dY = Raw(B * H * W * Cout * 4)
    assert L.lfrv_op_conv3x3(B, H, W, Cin, Cout, dX.ptr, dW.ptr, db.ptr, dY.ptr, None) == 0
    lfrv.device_sync()
    Y = dY.get(np.float32, (B, H, W, Cout))
    want = oracle.conv3x3(from_bf16(X), from_bf16(Wt), b.astype(np.float64))
    assert np.max(np.abs(Y - want)) <= 2e-5 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("B,H,W,Cin,Cout", [(1, 6, 10, 32, 8), (2, 17, 9, 64, 32), (1, 32, 32, 32, 64)])
def test_conv3x3_wgrad_matches_oracle(B, H, W, Cin, Cout):
    """dW[co][t][ci] = sum_p dU[p][co] X[p + off(t)][ci]: for each tap t the
    shifted, zero-padded input against the output gradient."""
    print(f"conv3x3 wgrad B={B} {H}x{W} {Cin}->{Cout}", flush=True)
    L = cnn_lib()
    X = to_bf16(RNG.standard_normal((B, H, W, Cin)).astype(np.float32))
    G = to_bf16(RNG.standard_normal((B, H, W, Cout)).astype(np.float32))
    dX, dG = Raw(X), Raw(G)
    dWb = Raw(Cout * 9 * Cin * 4)
    assert L.lfrv_op_conv3x3_wgrad(B, H, W, Cin, Cout, dG.ptr, dX.ptr, dWb.ptr, None) == 0
    lfrv.device_sync()
    got = dWb.get(np.float32, (Cout, 9, Cin))
    Xf, Gf = from_bf16(X).astype(np.float64), from_bf16(G).astype(np.float64)
    Xp = np.pad(Xf, ((0, 0), (1, 1), (1, 1), (0, 0)))
    want = np.zeros((Cout, 9, Cin))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        Xs = Xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]
        want[:, t, :] = Gf.reshape(-1, Cout).T @ Xs.reshape(-1, Cin)
    assert np.max(np.abs(got - want)) <= 2e-5 * max(1.0, np.abs(want).max())


def _random_params(levels, base, cpad, in_channels, scale=1.0):
    n = oracle.cnn_num_params(levels, base, cpad)
    w = np.zeros(n, np.float32)
    for name, ws, bs, ci, co in layer_slices(levels, base, cpad):
        real_ci = in_channels if name == "enc0a" else ci
        real_co = 3 if name == "head" else co
        Wl = np.zeros((co, 9, ci), np.float32)
        Wl[:real_co, :, :real_ci] = scale * RNG.standard_normal((real_co, 9, real_ci)) * np.sqrt(2.0 / (9 * real_ci))
        w[ws] = Wl.ravel()
        w[bs][:real_co] = 0.05 * RNG.standard_normal(real_co)
        bb = np.zeros(co, np.float32)
        bb[:real_co] = 0.05 * RNG.standard_normal(real_co)
        w[bs] = bb
    return w


@pytest.mark.parametrize("levels,in_channels,B,H,W", [(2, 32, 2, 8, 8), (4, 100, 1, 16, 24)])
def test_network_matches_oracle(levels, in_channels, B, H, W):
    print(f"U-Net levels={levels} in={in_channels} B={B} {H}x{W}", flush=True)
    net = Cnn(in_channels, levels=levels, base=32, max_batch=B, max_h=H, max_w=W, seed=3)
    cpad = net.cpad
    assert net.nparams() == oracle.cnn_num_params(levels, 32, cpad)
    w = _random_params(levels, 32, cpad, in_channels)
    net.set_params(w)
    x = RNG.uniform(0, 1, (B, H, W, in_channels)).astype(np.float32)
    base = RNG.uniform(0, 1, (B, 3, H, W)).astype(np.float32)
    target = (base + 0.1 * RNG.standard_normal((B, 3, H, W))).astype(np.float32)
    net.set_batch_host(x, base, target)
    pred = net.forward()
    xp = np.zeros((B, H, W, cpad))
    xp[..., :in_channels] = x
    wb = from_bf16(to_bf16(w)).astype(np.float64)
    loss_o, pred_o, g_o = oracle.cnn_forward_backward(levels, 32, xp, wb, base, target, round_bf16=True)
    d_o = pred_o - base
    assert np.max(np.abs((pred - base) - d_o)) <= 2e-2 * np.abs(d_o).max() + 1e-5
    loss = net.loss_grad()
    assert abs(loss - loss_o) <= 1e-3 * loss_o
    g = net.get_grads()
    # bf16 storage: where an FP32 sum and the oracle's FP64 sum round to
    # different bf16 values, the difference travels down the layers.  How far
    # is the oracle's own spread: the same oracle on the input scaled by
    # 1 + 1e-6 noise (far below a bf16 step) moves a deep layer's gradient by
    # several percent.  Each layer's error against the oracle, and the whole
    # gradient's, stays within twice that spread (at least 1e-2), and the
    # gradient points the same way.
    spread = {"all": 0.0}
    rng = np.random.default_rng(11)
    for _ in range(4):
        xq = xp * (1.0 + 1e-6 * rng.standard_normal(xp.shape))
        _, _, g_q = oracle.cnn_forward_backward(levels, 32, xq, wb, base, target, round_bf16=True)
        spread["all"] = max(spread["all"], rel_err(g_q, g_o))
        for name, ws, bs, ci, co in layer_slices(levels, 32, cpad):
            spread[name] = max(spread.get(name, 0.0), rel_err(g_q[ws], g_o[ws]), rel_err(g_q[bs], g_o[bs]))
    for name, ws, bs, ci, co in layer_slices(levels, 32, cpad):
        e_w = rel_err(g[ws], g_o[ws])
        e_b = rel_err(g[bs], g_o[bs])
        print(name, "rel err W %.2e b %.2e (oracle spread %.2e)" % (e_w, e_b, spread[name]))
        assert max(e_w, e_b) <= max(2.0 * spread[name], 1e-2), name
    print("whole gradient rel err %.2e (oracle spread %.2e)" % (rel_err(g, g_o), spread["all"]))
    assert rel_err(g, g_o) <= max(2.0 * spread["all"], 1e-2)
    cos = float(np.dot(g, g_o) / (np.linalg.norm(g) * np.linalg.norm(g_o)))
    assert cos > 0.995
    net.close()


def test_training_lowers_the_loss():
    """Fit base + a fixed smooth function of the input channels: 80 Adam steps
    lower the loss by 5x or more."""
    levels, cin, B, H, W = 3, 32, 2, 16, 16
    net = Cnn(cin, levels=levels, base=32, max_batch=B, max_h=H, max_w=W, seed=5)
    x = RNG.uniform(0, 1, (B, H, W, cin)).astype(np.float32)
    base = RNG.uniform(0, 1, (B, 3, H, W)).astype(np.float32)
    target = base + 0.3 * np.moveaxis(x[..., :3] - 0.5, -1, 1)
    net.set_batch_host(x, base, target.astype(np.float32))
    first = net.train_step(1e-3)
    for _ in range(79):
        last = net.train_step(1e-3)
    print("loss", first, "->", last)
    assert last < first / 5
    net.close()


@pytest.mark.parametrize("preset,dsf", [(0, 1), (0, 2), (0, 4), (3, 4)])
def test_warped_stack_matches_the_splat(preset, dsf):
    """The stack (anchor splat + box sum per group of 4 views) against the
    oracle's footprint splat of each view, normalised; channels beyond 4 V
    zero.  9 views (groups of 4, 4, 1) or the 25-view generator."""
    d = lfrv.synth_desc(preset, 3)
    d.dsf = d.render_scale = dsf
    if preset == 3:
        d.w = d.h = 32
        d.focal, d.baseline = 32.0, 10.0 / 32
    s = lfrv.synth(d)
    print(f"warped stack dsf={s['dsf']} V={s['V']}", flush=True)
    V, h, w, H, W = s["V"], s["h"], s["w"], s["H"], s["W"]
    L = cnn_lib()
    cpad = (4 * V + 31) // 32 * 32
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.warps_from_depth()
        ctx.prepare()
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        out = Raw(H * W * cpad * 2)
        assert L.lfrv_warped_stack(ctx._h, out.ptr, cpad) == 0
        ctx.sync()
        st = from_bf16(out.get(np.uint16, (H, W, cpad)))
    assert np.all(st[..., 4 * V:] == 0)
    for v in range(V):
        acc = np.zeros((3, H, W))
        wsum = np.zeros((H, W))
        oracle.splat(W, H, s["dsf"], tx[v], ty[v], s["views"][v].astype(np.float64), out=acc, wsum=wsum)
        col = np.where(wsum > 0, acc / np.where(wsum > 0, wsum, 1), 0.0)
        for c in range(3):
            assert np.max(np.abs(st[..., 4 * v + c] - col[c])) <= 8e-3 * max(1.0, col.max())
        assert np.max(np.abs(st[..., 4 * v + 3] - np.minimum(wsum, 1.0))) <= 8e-3


def test_warm_start_keeps_the_lower_energy_start():
    """An untrained network's start and the splatted average: the warm start
    keeps whichever has the lower normal-equation energy, and the solve from
    it reaches the same solution as the plain solve."""
    s = lfrv.synth(0, seed=3)
    V, H, W = s["V"], s["H"], s["W"]
    L = cnn_lib()
    with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.ibr_gradient(1e-6, 5000)
        U_plain = ctx.get_solution()
    for scale in (0.0, 30.0):       # head zero: the same start; a large random head: a worse one
        net = Cnn(4 * V, levels=3, base=32, max_batch=1, max_h=H, max_w=W, seed=9)
        w = net.get_params()
        head_ws, head_bs = layer_slices(3, 32, net.cpad)[-1][1:3]
        w[head_ws] *= scale
        w[head_bs] = 0.0
        net.set_params(w)
        with lfrv.Context(pipelines.gradient_ibr_params(s)) as ctx:
            pipelines.load_scene(ctx, s)
            ctx.clear_solution()
            ctx.warps_from_depth()
            ctx.prepare()
            ctx.setup_visibility()
            ctx.compute_weights()
            ctx.compute_initial_image()
            ctx.init_u_gradient()
            U0 = ctx.get_solution()
            e0 = ctx.gibr_energy()
            ws = WarmStats()
            assert L.lfrv_warm_start(ctx._h, net.h, ctypes.byref(ws)) == 0
            assert abs(ws.energy_plain - e0) <= 1e-9 * e0
            U1 = ctx.get_solution()
            if ws.used_net:
                assert ws.energy_net < ws.energy_plain
                assert abs(ctx.gibr_energy() - ws.energy_net) <= 1e-9 * ws.energy_net
            else:
                assert np.array_equal(U1, U0)
            if scale == 30.0:
                assert ws.used_net == 0 and ws.energy_net > ws.energy_plain
            st = ctx.cg_solve(1e-6, 5000)
            assert all(st["converged"])
            assert rel_err(ctx.get_solution(), U_plain) < 1e-4
        net.close()


def test_device_scene_renderer_matches_the_host_one():
    """lfrv_synth_generate_device draws the scene with the host code and runs the
    same FP64 ray tracer per pixel: views, depths and the target agree with
    lfrv_synth_generate to float rounding."""
    from lfrv.cnn import cnn_lib as _cl
    from lfrv.warmstart import SceneOnDevice, scene_desc
    _cl()
    for w, dsf in ((32, 2), (48, 4)):
        d = scene_desc(11, w, dsf)
        print(f"device renderer {w}x{w} dsf {dsf}", flush=True)
        host = lfrv.synth(d)
        dev = SceneOnDevice(d, with_target=True)
        V, n = host["V"], w * w
        views = dev.views.get(np.float32, (V, 3, w, w))
        depths = dev.depths.get(np.float32, (V, w, w))
        target = dev.target.get(np.float32, (3, w * dsf, w * dsf))
        assert np.max(np.abs(views - host["views"])) <= 1e-6
        assert np.max(np.abs(depths - host["depths"])) <= 1e-5 * np.abs(host["depths"]).max()
        assert np.max(np.abs(target - host["target"])) <= 1e-6
        np.testing.assert_array_equal(dev.P_views, host["P_views"])
        np.testing.assert_array_equal(dev.P_target, host["P_target"])
        dev.free()


def test_trained_warm_start_saves_iterations_on_held_out_scenes():
    """The learned warm start at a small scale: 400 scenes of the 5 x 5-view
    generator at its 4x upsampling (32^2 views, 128^2 targets) solved by the
    plain method to relative residual 1e-5, the U-Net trained on them for 6
    epochs, then 200 held-out scenes solved both ways.  Every warm-started
    solve reaches the same tolerance and the plain solution (both within the
    solver's tolerance of the exact solution), the network's start is used
    only where it lowers the energy, and on average it saves CG iterations."""
    from lfrv import warmstart as wsm
    print("warm start: training set of 400 scenes", flush=True)
    data = wsm.TrainingSet(list(range(1000, 1400)), 32)
    net = Cnn(data.in_channels, levels=4, base=32, max_batch=4, max_h=data.H, max_w=data.W, seed=7)
    losses = wsm.train(net, data, epochs=6, batch=4, lr=2e-3, log=print)
    assert losses[-1] < losses[0]
    print("warm start: 200 held-out scenes", flush=True)
    res = wsm.compare(net, list(range(5000, 5200)), 32)
    net.close()
    data.free()
    plain = np.array([r["iters_plain"] for r in res])
    warm = np.array([r["iters_warm"] for r in res])
    used = np.array([r["used_net"] for r in res])
    rd = np.array([r["rel_diff"] for r in res])
    print("iterations plain %.2f warm %.2f, net used in %d/200, max rel diff %.2e" %
          (plain.mean(), warm.mean(), used.sum(), rd.max()))
    for r in res:
        assert r["true_rel_residual_warm"] <= 1.5e-5
        assert r["rel_diff"] <= 1e-3
        if r["used_net"]:
            assert r["energy_net_start"] < r["energy_plain_start"]
    assert warm.mean() < plain.mean()
