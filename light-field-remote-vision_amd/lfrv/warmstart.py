"""The learned warm start end to end over the C-ABI: training sets made by
the solver itself on the GPU, training of the U-Net on the GPU, and the
comparison of warm-started and plain solves.

Scenes come from the 5 x 5-view generator (8 random tilted planes, disparity
-2 .. +2 px per grid step, views box-averaged from 4x renders), rendered on
the GPU (lfrv_synth_generate_device), with that generator's 4x upsampling
(dsf 4).  The solver is the gradient-constraint IBR (IBR_gradient's normal
equations, lfrv_cg_solve) to relative residual 1e-5; the network's input is
the stack of the 25 warped views with their validity (100 channels) at the
target resolution, its base the solver's own starting image (the splatted
average with hole filling), its target the converged solution.  Everything
stays in device memory.  (At dsf 2 the plain PCG converges in about 9
iterations from the splatted average, which leaves a start nothing to save;
at dsf 4 it takes about 35.)
"""
import ctypes
import time

import numpy as np

from . import Context, LfrvError, synth_desc
from .cnn import Cnn, Raw, WarmStats, cnn_lib
from .pipelines import gradient_ibr_params

TOL = 1e-5
MAXIT = 5000
DSF = 4


def scene_desc(seed, w, dsf=DSF):
    """The 5 x 5-view generator at w x w views and a (w dsf)^2 target: the
    focal length is w and the baseline 10 / w, so the disparities in view
    pixels are the same at every w."""
    d = synth_desc(3, seed)
    d.w = d.h = w
    d.dsf = dsf
    d.focal = float(w)
    d.baseline = 10.0 / w
    return d


class SceneOnDevice:
    """One generated scene: views and depths in device memory, cameras on the host."""

    def __init__(self, desc, stream=None, with_target=False):
        L = cnn_lib()
        self.V = desc.grid * desc.grid
        self.w, self.h, self.dsf = desc.w, desc.h, desc.dsf
        self.W, self.H = self.w * self.dsf, self.h * self.dsf
        n = self.w * self.h
        self.views = Raw(self.V * 3 * n * 4)
        self.depths = Raw(self.V * n * 4)
        self.target = Raw(3 * self.W * self.H * 4) if with_target else None
        self.P_views = np.zeros((self.V, 3, 4), np.float64)
        self.P_target = np.zeros((3, 4), np.float64)
        st = L.lfrv_synth_generate_device(ctypes.byref(desc), self.views.ptr, self.depths.ptr,
                                          self.P_views.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          self.P_target.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          self.target.ptr if self.target else None, None, stream)
        if st != 0:
            raise LfrvError(st, "lfrv_synth_generate_device failed")

    def load(self, ctx):
        n = self.w * self.h
        for i in range(self.V):
            ctx._check(ctx_lib().lfrv_set_view(ctx._h, i, ctypes.c_void_p(self.views.ptr.value + i * 3 * n * 4)))
            Pd = np.ascontiguousarray(self.P_views[i], np.float64).reshape(12)
            ctx._check(ctx_lib().lfrv_set_depth(ctx._h, i, ctypes.c_void_p(self.depths.ptr.value + i * n * 4),
                                                Pd.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        ctx.set_target_camera(self.P_target)

    def free(self):
        for b in (self.views, self.depths, self.target):
            if b is not None:
                b.free()


def ctx_lib():
    from . import lib
    return lib()


def solver_start(ctx):
    """IBR_gradient up to the CG (gradientIBR.cpp:12-150): warps, splat
    structure, visibility, weights, initial image with hole filling, target
    gradient; U holds the splatted-average start."""
    ctx.clear_solution()
    ctx.warps_from_depth()
    ctx.prepare()
    ctx.setup_visibility()
    ctx.compute_weights()
    ctx.compute_initial_image()
    ctx.init_u_gradient()


class TrainingSet:
    """n solved scenes in device memory: input stack (bf16 NHWC, cpad
    channels), start image and converged solution (planar FP32) per scene,
    and the plain solve's CG iterations."""

    def __init__(self, seeds, w, dsf=DSF, log=None):
        d0 = scene_desc(seeds[0], w, dsf)
        self.V = d0.grid * d0.grid
        self.in_channels = 4 * self.V
        self.cpad = (self.in_channels + 31) // 32 * 32
        self.W = self.H = w * dsf
        N = self.W * self.H
        self.n = len(seeds)
        self.x = Raw(self.n * N * self.cpad * 2)
        self.u0 = Raw(self.n * 3 * N * 4)
        self.ustar = Raw(self.n * 3 * N * 4)
        self.iters = np.zeros(self.n, np.int64)
        L = cnn_lib()
        t0 = time.perf_counter()
        with Context(gradient_ibr_params(dict(V=self.V, w=w, h=w, dsf=dsf))) as ctx:
            for k, seed in enumerate(seeds):
                sc = SceneOnDevice(scene_desc(seed, w, dsf))
                sc.load(ctx)
                solver_start(ctx)
                ctx._check(L.lfrv_warped_stack(ctx._h, self.x_ptr(k), self.cpad))
                ctx._check(ctx_lib().lfrv_get_solution(ctx._h, self.u0_ptr(k)))
                st = ctx.cg_solve(TOL, MAXIT)
                ctx._check(ctx_lib().lfrv_get_solution(ctx._h, self.ustar_ptr(k)))
                self.iters[k] = max(st["iterations"])
                sc.free()
                if log and (k + 1) % 250 == 0:
                    log(f"training set: {k + 1}/{self.n} scenes solved, {time.perf_counter() - t0:.1f} s")
        self.seconds = time.perf_counter() - t0

    def x_ptr(self, k):
        return ctypes.c_void_p(self.x.ptr.value + int(k) * self.W * self.H * self.cpad * 2)

    def u0_ptr(self, k):
        return ctypes.c_void_p(self.u0.ptr.value + int(k) * 3 * self.W * self.H * 4)

    def ustar_ptr(self, k):
        return ctypes.c_void_p(self.ustar.ptr.value + int(k) * 3 * self.W * self.H * 4)

    def free(self):
        for b in (self.x, self.u0, self.ustar):
            b.free()


def train(net, data, epochs, batch, lr, seed=0, log=None):
    """Adam over the training set, shuffled each epoch; the learning rate
    decays by cos to 5 % of lr.  Returns the per-epoch mean losses."""
    rng = np.random.default_rng(seed)
    steps_per_epoch = data.n // batch
    total = epochs * steps_per_epoch
    losses = []
    step = 0
    t0 = time.perf_counter()
    for e in range(epochs):
        order = rng.permutation(data.n)
        acc = 0.0
        for s in range(steps_per_epoch):
            idx = order[s * batch:(s + 1) * batch]
            net.set_batch(batch, data.H, data.W, [data.x_ptr(k) for k in idx], [data.u0_ptr(k) for k in idx],
                          [data.ustar_ptr(k) for k in idx])
            f = 0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * step / max(total, 1)))
            acc += net.train_step(lr * f)
            step += 1
        losses.append(acc / max(steps_per_epoch, 1))
        if log:
            log(f"epoch {e + 1}/{epochs}: mean loss {losses[-1]:.3e} ({time.perf_counter() - t0:.1f} s)")
    return losses


def compare(net, seeds, w, dsf=DSF):
    """Plain and warm-started solves of the same scenes.  Per scene: CG
    iterations (max over channels) of both, whether the network's start was
    kept, the energies of both starts, the relative difference of the two
    solutions, and the device-resident wall time of each form (from the
    views, depths and cameras in memory to the solution)."""
    L = cnn_lib()
    out = []
    V = 25
    with Context(gradient_ibr_params(dict(V=V, w=w, h=w, dsf=dsf))) as ctx:
        for seed in seeds:
            sc = SceneOnDevice(scene_desc(seed, w, dsf))
            sc.load(ctx)
            # plain
            ctx.sync()
            t0 = time.perf_counter()
            solver_start(ctx)
            st_p = ctx.cg_solve(TOL, MAXIT)
            ctx.sync()
            ms_p = 1e3 * (time.perf_counter() - t0)
            U_p = ctx.get_solution()
            # warm
            ctx.sync()
            t0 = time.perf_counter()
            solver_start(ctx)
            ws = WarmStats()
            ctx._check(L.lfrv_warm_start(ctx._h, net.h, ctypes.byref(ws)))
            st_w = ctx.cg_solve(TOL, MAXIT)
            ctx.sync()
            ms_w = 1e3 * (time.perf_counter() - t0)
            U_w = ctx.get_solution()
            out.append(dict(seed=seed, iters_plain=max(st_p["iterations"]), iters_warm=max(st_w["iterations"]),
                            used_net=ws.used_net, energy_plain_start=ws.energy_plain, energy_net_start=ws.energy_net,
                            rel_diff=float(np.linalg.norm(U_w - U_p) / np.linalg.norm(U_p)),
                            true_rel_residual_warm=max(st_w["true_rel_residual"]), ms_plain=ms_p, ms_warm=ms_w))
            sc.free()
    return out
