"""Batches of solves (lfrv_batch) against the same solves made one at a time:
PCG solves of uneven sizes and view counts iterated together through one
captured graph, and the IBR_color solves' steps batched one launch per step
beside them."""
import numpy as np
import pytest

import lfrv
from lfrv import pipelines
from lfrv.batch import Batch

pytestmark = pytest.mark.gpu


def cg_scene(seed, W, H, grid):
    d = lfrv.synth_desc(2, seed)
    d.grid = grid
    d.w, d.h = W, H
    d.focal = 1000.0 * W / 1024.0
    return lfrv.synth(d)


def pd_scene(seed, w, grid=3):
    d = lfrv.synth_desc(1, seed)
    d.grid = grid
    d.w = d.h = w
    d.dsf = d.render_scale = 2
    d.focal = float(w)
    d.baseline = 0.7 * 4.0 / (2 * w)
    return lfrv.synth(d)


CG_CASES = [(11, 64, 64, 2), (12, 96, 80, 3), (13, 130, 70, 4), (14, 72, 128, 5), (15, 200, 136, 3),
            (16, 64, 48, 2)]
PD_CASES = [(21, 40, 20), (22, 56, 35), (23, 32, 10)]


def _cg_ctx(s):
    ctx = lfrv.Context(pipelines.gradient_ibr_params(s))
    pipelines.load_scene(ctx, s)
    return ctx


def _pd_ctx(s):
    prm = pipelines.color_sr_params(s, lambda_=0.05, lambda_max_factor=100.0, inner_iterations=1, dt_beta=0.3,
                                    gradient_step=1.0)
    ctx = lfrv.Context(prm)
    pipelines.load_scene(ctx, s)
    return ctx


@pytest.mark.parametrize("use_graph,chunk", [(True, 8), (False, 4), (True, 2)])
def test_batch_equals_single_solves(use_graph, chunk):
    print(f"batch of {len(CG_CASES)} CG + {len(PD_CASES)} PD solves, graph={use_graph} chunk={chunk}", flush=True)
    cg_scenes = [cg_scene(*c) for c in CG_CASES]
    pd_scenes = [pd_scene(*c[:2]) for c in PD_CASES]
    # one at a time
    want = []
    for s in cg_scenes:
        with _cg_ctx(s) as ctx:
            st = ctx.ibr_gradient(1e-5, 5000)
            want.append((ctx.get_solution(), max(st["iterations"]), None))
    for s, c in zip(pd_scenes, PD_CASES):
        with _pd_ctx(s) as ctx:
            ctx.clear_solution()
            pipelines.color_sr_setup(ctx)
            ctx.sr_init_regularizer_weight()
            ctx.sr_iterate(c[2])
            want.append((ctx.get_solution(), c[2], (ctx.get_buffer("UQ"), ctx.get_buffer("NORM_MASK"))))
    # together, interleaved kinds, run twice (the second run replays the same graph)
    ctxs = []
    b = Batch(chunk=chunk, use_graph=use_graph)
    order = []
    for k in range(max(len(cg_scenes), len(pd_scenes))):
        if k < len(cg_scenes):
            ctx = _cg_ctx(cg_scenes[k])
            ctxs.append(ctx)
            b.add_cg(ctx, 1e-5, 5000)
            order.append(k)
        if k < len(pd_scenes):
            ctx = _pd_ctx(pd_scenes[k])
            ctxs.append(ctx)
            b.add_pd(ctx, PD_CASES[k][2])
            order.append(len(cg_scenes) + k)
    for rep in range(2):
        res = b.run()
        for j, (ctx, r) in enumerate(zip(ctxs, res)):
            U, iters, pd_out = want[order[j]]
            assert r["status"] == lfrv.OK
            assert r["iterations"] == iters, (j, r, iters)
            if r["kind"] == 0:
                assert r["true_rel_residual"] <= 1e-5
            assert np.array_equal(ctx.get_solution(), U), (rep, j)
            if pd_out is not None:  # the over-relaxed solution and the normalisation buffer too
                assert np.array_equal(ctx.get_buffer("UQ"), pd_out[0]), (rep, j)
                assert np.array_equal(ctx.get_buffer("NORM_MASK"), pd_out[1]), (rep, j)
            assert r["done_ms"] > 0
    b.close()
    for ctx in ctxs:
        ctx.close()


def test_batch_rejects_what_it_cannot_batch():
    s = cg_scene(31, 48, 48, 2)
    d = lfrv.synth_desc(2, 32)
    d.grid, d.w, d.h, d.dsf, d.render_scale, d.focal = 2, 32, 32, 2, 2, 2000.0 * 32 / 1024
    s2 = lfrv.synth(d)
    b = Batch()
    with _cg_ctx(s2) as ctx:          # dsf 2: not a stencil PCG solve
        with pytest.raises(lfrv.LfrvError) as ei:
            b.add_cg(ctx)
        assert ei.value.status == lfrv.E_ARG
    with _cg_ctx(s) as ctx:
        with pytest.raises(lfrv.LfrvError):
            b.add_pd(ctx, 10)             # an IBR_gradient context is not a PD solve
    b.close()
