"""The IBR_color driver loop (super-resolution.cpp:177-300) run by the FP64
oracle against lfrv_sr_solve on the GPU: the oracle's FISTA steps, energies
and moving-pixel counts decide with the driver's float rules (EPSILON_TOL
1e-4, MAX_DIFF_THRESHOLD 1e-3, PERCENT_MOVING 0.01), the GPU loop decides on
the device.  Both must stop their passes at the same steps for the same
reasons, and end with energies within 1e-4 relative and solutions within
FP32 accumulation error.  The cases are chosen with a margin on every
rule, so FP32 and FP64 decide alike."""
import numpy as np
import pytest

import lfrv
import oracle
from lfrv import pipelines

pytestmark = pytest.mark.gpu

EPS, MAXD, PCT = np.float32(1e-4), 1e-3, np.float32(0.01)


def _moving(A, B):
    """sr_compare_images (super-resolution.cpp:16-51): pixels whose largest
    channel change is >= MAX_DIFF_THRESHOLD."""
    return int((np.max(np.abs(np.asarray(A, np.float64) - np.asarray(B, np.float64)), axis=0) >= MAXD).sum())


class OracleColorSR:
    """sr_synthesize_view's state after FINISHED INIT, in FP64, on given warps."""

    def __init__(self, s, tx, ty, dpx, dpy, prm):
        self.s, self.tx, self.ty, self.dpx, self.dpy = s, tx, ty, dpx, dpy
        self.W, self.H, self.dsf = s["W"], s["H"], s["dsf"]
        self.sig = float(prm.sigma_sensor)
        self.lam, self.lmf = float(prm.lambda_), float(prm.lambda_max_factor)
        self.alpha, self.beta = float(prm.dt_alpha), float(prm.dt_beta)
        self.gstep, self.gw, self.inner = float(prm.gradient_step), prm.gw_type, prm.inner_iterations
        W, H, dsf = self.W, self.H, self.dsf
        self.st = oracle.SrState(W, H, np.zeros((3, H, W)))
        st = self.st
        om, gx, gy = oracle.sr_weights(W, H, dsf, tx, ty, dpx, dpy, st.U, np.zeros((H, W), np.uint8), st.PX[0],
                                       st.PY[0], self.sig, 1.0)
        st.PX[0], st.PY[0] = gx, gy
        cov = oracle.coverage(W, H, dsf, tx, ty)
        self.vis = (cov >= 1.0).astype(np.uint8)
        st.norm_mask += cov
        U0, st.norm_mask = oracle.initial_image(W, H, dsf, tx, ty, s["views"], om)
        st.U = oracle.push_pull(U0, self.vis)
        st.Uq = st.U.copy()
        self.om, self.g = om, None

    def compute_weights(self):
        st = self.st
        self.om, gx, gy = oracle.sr_weights(self.W, self.H, self.dsf, self.tx, self.ty, self.dpx, self.dpy, st.U,
                                            self.vis, st.PX[0], st.PY[0], self.sig, 1.0)
        st.PX[0], st.PY[0] = gx, gy

    def init_regularizer_weight(self):
        self.g, self.st.norm_mask = oracle.regularizer_weight(self.W, self.H, self.dsf, self.tx, self.ty, self.om,
                                                              self.sig, self.lmf * self.lam, self.lam,
                                                              self.st.norm_mask)

    def energy(self):
        e, _ = oracle.sr_energy(self.W, self.H, self.dsf, self.tx, self.ty, self.s["views"], self.om, self.sig,
                                self.lam, self.alpha, self.beta, self.gw, self.vis, self.st.U)
        return e

    def step(self):
        oracle.fista_step(self.W, self.H, self.dsf, self.tx, self.ty, self.s["views"], self.om, self.sig, self.alpha,
                          self.beta, self.gstep, self.gw, self.lam, self.inner, self.g, self.st)


def oracle_driver(o, niter, nmeta):
    """super-resolution.cpp:177-300 over the oracle, the rules in float (cuflt).
    Returns what the loop did and the margin of every decision it took (how
    far the tested quantity was from its threshold, relative)."""
    f = np.float32
    WH = o.W * o.H
    prev_meta = o.st.U.copy()
    prev = prev_meta.copy()
    energy = prev_e = prev_e_rw = f(0)
    info = dict(steps=0, meta=0, stops=[], converged=0, margins=[])
    for r in range(nmeta):
        info["meta"] = r + 1
        o.compute_weights()
        o.init_regularizer_weight()
        energy = f(o.energy())
        prev_e = energy
        stop = 0
        for k in range(niter):
            o.step()
            info["steps"] += 1
            energy = f(o.energy())
            info["margins"].append(abs(float(energy) / 1e-4 - 1))
            if energy < EPS:
                stop = 1
                break
            rel = abs((energy - prev_e) / energy)
            info["margins"].append(abs(float(rel) / 1e-4 - 1))
            if rel < EPS:
                stop = 2
                break
            prev_e = energy
            frac = f(_moving(o.st.U, prev)) / f(WH)
            info["margins"].append(abs(float(frac) / 0.01 - 1))
            if frac < PCT:
                stop = 3
                break
            prev = o.st.U.copy()
        info["stops"].append(stop)
        if energy < EPS:
            if prev_e_rw < EPS:
                info["converged"] = 1
                break
            prev_e_rw = energy
            continue
        frac = f(_moving(o.st.U, prev_meta)) / f(WH)
        info["margins"].append(abs(float(frac) / 0.01 - 1))
        if frac < PCT:
            info["converged"] = 1
            break
        prev_meta = o.st.U.copy()
    info["energy"] = energy
    return info


def color_scene(dsf, w=40, flat=False, seed=5):
    d = lfrv.synth_desc(1, seed)
    d.w = d.h = w
    d.dsf = d.render_scale = dsf
    d.focal = float(w)
    d.baseline = 0.7 * 4.0 / (w * dsf)
    s = lfrv.synth(d)
    if flat:  # one colour everywhere: TV and gradient terms vanish
        s["views"][:] = np.array([0.2, 0.5, 0.7], np.float32)[None, :, None, None]
    return s


# (dsf, inner, beta, niter, nmeta, flat, scale, seed): every pass to niter;
# the moving rule, then the relative-change rule, on long passes (every
# decision of the oracle's loop at least 25 % away from its threshold); the
# small-energy rules
CASES = [
    (2, 3, 0.3, 15, 3, False, 1.0, 5),
    (2, 2, 0.0, 200, 2, False, 1.0, 7),
    (2, 1, 0.3, 20, 3, True, 1e-9, 5),
]


@pytest.mark.parametrize("dsf,inner,beta,niter,nmeta,flat,scale,seed", CASES)
def test_sr_solve_matches_the_oracle_driver(dsf, inner, beta, niter, nmeta, flat, scale, seed):
    s = color_scene(dsf, flat=flat, seed=seed)
    prm = pipelines.color_sr_params(s, lambda_=0.05 * scale, lambda_max_factor=100.0, inner_iterations=inner,
                                    dt_alpha=scale, dt_beta=beta * scale, gradient_step=1.0,
                                    sigma_sensor=np.float32(1.0) / np.float32(255))
    V, h, w = s["V"], s["h"], s["w"]
    print(f"sr loop vs oracle dsf={dsf} inner={inner} niter={niter} nmeta={nmeta} flat={flat}", flush=True)
    with lfrv.Context(prm) as ctx:
        pipelines.load_scene(ctx, s)
        ctx.clear_solution()
        pipelines.color_sr_setup(ctx)
        tx = ctx.get_buffer("TAU_X").reshape(V, h, w)
        ty = ctx.get_buffer("TAU_Y").reshape(V, h, w)
        dpx = ctx.get_buffer("DPART_X").reshape(V, h, w)
        dpy = ctx.get_buffer("DPART_Y").reshape(V, h, w)
        st = ctx.sr_solve(niter, nmeta)
        U = ctx.get_solution()
    o = OracleColorSR(s, tx, ty, dpx, dpy, prm)
    info = oracle_driver(o, niter, nmeta)
    print(info["stops"], info["steps"], st, min(info["margins"]))
    assert min(info["margins"]) > 0.2  # no decision of the oracle's loop sits on its threshold
    assert st["fista_steps"] == info["steps"]
    assert st["meta_iterations"] == info["meta"]
    assert st["last_stop"] == (info["stops"][-1] if info["stops"] else 0)
    assert st["converged"] == info["converged"]
    assert abs(float(st["energy"]) - float(info["energy"])) <= 1e-4 * abs(float(info["energy"]))
    assert np.max(np.abs(U - o.st.U)) < 2e-3
