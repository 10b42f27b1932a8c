"""ctypes bindings of the FP64 oracle (oracle/lfrv_oracle.c, oracle/lfrv_oracle_cnn.c).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module.  It is the checker of the HIP path,
never part of it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblfrv_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_u8 = ctypes.POINTER(ctypes.c_ubyte)
_i = ctypes.c_int
_d = ctypes.c_double


class CgStats(ctypes.Structure):
    _fields_ = [("iterations", _i * 3), ("converged", _i * 3), ("rel_residual", _d * 3),
                ("true_rel_residual", _d * 3), ("status", _i)]


_lib = None


def build():
    subprocess.check_call(["make", "-B", "-C", _HERE], stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_footprint": (_i, [ctypes.c_float, ctypes.c_float, _i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), _dp, _dp]),
        "orc_cell_valid": (_i, [ctypes.c_float, ctypes.c_float, _i, _i]),
        "orc_gather": (None, [_i, _i, _i, _i, _i, _fp, _fp, _dp, _i, _dp]),
        "orc_splat": (None, [_i, _i, _i, _i, _i, _fp, _fp, _dp, _i, _dp, _dp, _dp]),
        "orc_target_visibility": (None, [_i, _i, _i, _i, _i, _fp, _fp, _u8]),
        "orc_sobel": (None, [_i, _i, _dp, _dp, _dp, _dp, _dp, _u8]),
        "orc_laplacian": (None, [_i, _i, _dp, _dp, _u8]),
        "orc_dtd": (None, [_i, _i, _dp, _dp]),
        "orc_dtg": (None, [_i, _i, _dp, _dp, _dp]),
        "orc_angular_weights": (None, [_i, _i, _dp, _dp, _d, _fp, _fp, _d, _dp]),
        "orc_deform_weights": (None, [_i, _i, _i, _i, _fp, _fp, _dp]),
        "orc_decompose_P": (None, [_dp, _dp, _dp, _dp, _dp]),
        "orc_warps_from_depth": (None, [_i, _i, _i, _i, _fp, _dp, _dp, _d, _i, _d, _fp, _fp, _fp, _fp, _fp]),
        "orc_gl_target_depth": (None, [_i, _i, _i, _i, _i, _fp, _fp, _fp, _dp, _dp, _i, _fp]),
        "orc_radial_to_ortho": (None, [_i, _i, _dp, _fp, _fp, _fp]),
        "orc_gl_warps": (None, [_i, _i, _i, _i, _i, _fp, _fp, _fp, _dp, _dp, _fp, _d, _fp, _fp, _fp, _fp, _fp]),
        "orc_initial_image": (None, [_i, _i, _i, _i, _i, _i, _fp, _fp, _fp, _dp, _dp, _dp]),
        "orc_hole_filling": (_i, [_i, _i, _dp, _u8, _i, _d]),
        "orc_u_gradient": (None, [_i, _i, _i, _i, _i, _i, _fp, _fp, _fp, _dp, _d, _dp, _dp, _dp, _dp]),
        "orc_apply_A": (None, [_i, _i, _i, _i, _i, _i, _fp, _fp, _dp, _d, _d, _d, _dp, _dp]),
        "orc_diag_A": (None, [_i, _i, _i, _i, _i, _i, _fp, _fp, _dp, _d, _d, _d, _dp]),
        "orc_rhs": (None, [_i, _i, _i, _i, _i, _i, _fp, _fp, _fp, _dp, _d, _d, _dp, _d, _dp, _dp, _dp]),
        "orc_pcg": (_i, [_i, _i, _i, _i, _i, _i, _fp, _fp, _dp, _d, _d, _d, _dp, _dp, _d, _i, _i, ctypes.POINTER(CgStats)]),
        "orc_gibr_energy": (_d, [_i, _i, _i, _i, _i, _i, _fp, _fp, _fp, _dp, _d, _d, _dp, _d, _dp, _dp, _dp]),
        "orc_pd_dual": (None, [_i, _i, _d, _dp, _dp, _dp, _dp]),
        "orc_pd_primal": (None, [_i, _i, _d, _d, _dp, _dp, _dp, _dp]),
        "orc_gauss_smooth": (None, [_i, _i, _i, _d, _dp, _dp]),
        "orc_regularizer_weight": (None, [_i, _i, _i, _i, _i, _i, _fp, _fp, _dp, _d, _d, _d, _dp, _dp]),
        "orc_fista_step": (None, [_i, _i, _i, _i, _i, _i, _fp, _fp, _fp, _dp, _d, _d, _d, _d, _i, _d, _i, _dp,
                                  _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_sr_energy": (_d, [_i, _i, _i, _i, _i, _i, _fp, _fp, _fp, _dp, _d, _d, _d, _d, _i, _u8, _dp, _dp]),
        "orc_push_pull": (None, [_i, _i, _i, _dp, _u8]),
        "orc_psnr": (_d, [_i, _i, _dp, _dp]),
        "orc_conv3x3": (None, [_i, _i, _i, _i, _i, _dp, _dp, _dp, _dp]),
        "orc_cnn_num_params": (ctypes.c_size_t, [_i, _i, _i]),
        "orc_cnn_forward_backward": (_d, [_i, _i, _i, _i, _i, _i, _dp, _dp, _dp, _dp, _i, _dp, _dp]),
        "orc_num_threads": (_i, []),
        "orc_set_num_threads": (None, [_i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _D(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _F(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _U8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def pd(a):
    return _p(a, _dp)


def pf(a):
    return _p(a, _fp)


def pu8(a):
    return _p(a, _u8)


def set_threads(t):
    lib().orc_set_num_threads(int(t))


def num_threads():
    return lib().orc_num_threads()


def footprint(tx, ty, W, H, dsf):
    ks = dsf + 1
    x0, y0 = _i(), _i()
    wx, wy = np.zeros(ks), np.zeros(ks)
    r = lib().orc_footprint(tx, ty, W, H, dsf, ctypes.byref(x0), ctypes.byref(y0), pd(wx), pd(wy))
    return r, x0.value, y0.value, wx, wy


def cell_valid(tx, ty, W, H):
    return bool(lib().orc_cell_valid(tx, ty, W, H))


def gather(W, H, dsf, tx, ty, u):
    """u: (C, H, W) -> (C, h, w)."""
    tx, ty = _F(tx), _F(ty)
    h, w = tx.shape
    u = _D(u)
    C = u.shape[0]
    out = np.zeros((C, h, w))
    lib().orc_gather(W, H, w, h, dsf, pf(tx), pf(ty), pd(u), C, pd(out))
    return out


def structured_warp_down(u, dmap, dx, dy, vis, dsf):
    """The structured (disparity) model of cocolib's vtv_sr (not on the ULF
    path; the cross-check of the tau-footprint model for pure translations):
    vtv_sr_warp_view_device (vtv_sr_kernels.cu:72-135) samples u (C, H, W)
    bilinearly at (x + d dx, y + d dy), integer-origin pixels, renormalised
    where the right / bottom taps leave the image and 0 where the sample starts
    outside or the pixel is not visible; vtv_sr_downsample_view_device
    (:767-802) then averages the visible pixels of each dsf x dsf block.
    dmap, vis: (H, W) on the hi-res view grid.  Returns (C, H / dsf, W / dsf)."""
    u = _D(u)
    C, H, W = u.shape
    d = _D(dmap)
    vis = np.asarray(vis).astype(bool)
    oy, ox = np.mgrid[0:H, 0:W].astype(np.float64)
    ux = ox + d * dx
    uy = oy + d * dy
    px = np.floor(ux).astype(np.int64)
    py = np.floor(uy).astype(np.int64)
    ax, ay = ux - px, uy - py
    out_img = (px < 0) | (py < 0) | (px > W - 1) | (py > H - 1)
    pxc, pyc = np.clip(px, 0, W - 1), np.clip(py, 0, H - 1)
    xr, yr = px < W - 1, py < H - 1
    pxr, pyr = np.clip(px + 1, 0, W - 1), np.clip(py + 1, 0, H - 1)
    wmm, wpm, wmp, wpp = (1 - ax) * (1 - ay), ax * (1 - ay), (1 - ax) * ay, ax * ay
    weight = 1.0 - np.where(xr, 0.0, wpm) - np.where(yr, 0.0, wmp) - np.where(xr & yr, 0.0, wpp)
    v = np.zeros((C, H, W))
    for c in range(C):
        r = u[c][pyc, pxc] * wmm
        r = r + np.where(xr, u[c][pyc, pxr] * wpm, 0.0)
        r = r + np.where(yr, u[c][pyr, pxc] * wmp, 0.0)
        r = r + np.where(xr & yr, u[c][pyr, pxr] * wpp, 0.0)
        v[c] = np.where(vis & ~out_img, r / weight, 0.0)
    h, w = H // dsf, W // dsf
    vv = (v[:, :h * dsf, :w * dsf] * vis[None, :h * dsf, :w * dsf]).reshape(C, h, dsf, w, dsf).sum(axis=(2, 4))
    cnt = vis[:h * dsf, :w * dsf].reshape(h, dsf, w, dsf).sum(axis=(1, 3))
    return np.where(cnt > 0, vv / np.maximum(cnt, 1), 0.0)


def splat(W, H, dsf, tx, ty, vals, wgt=None, out=None, wsum=None):
    """vals: (C, h, w); accumulates into out (C, H, W) and wsum (H, W)."""
    tx, ty = _F(tx), _F(ty)
    h, w = tx.shape
    vals = _D(vals) if vals is not None else np.zeros((1, h, w))
    C = vals.shape[0]
    if out is None:
        out = np.zeros((C, H, W))
    if wsum is None:
        wsum = np.zeros((H, W))
    wg = None if wgt is None else _D(wgt)
    lib().orc_splat(W, H, w, h, dsf, pf(tx), pf(ty), pd(vals), C, pd(wg), pd(out), pd(wsum))
    return out, wsum


def target_visibility(W, H, dsf, taus):
    vis = np.zeros((H, W), np.uint8)
    for tx, ty in taus:
        tx, ty = _F(tx), _F(ty)
        h, w = tx.shape
        lib().orc_target_visibility(W, H, w, h, dsf, pf(tx), pf(ty), pu8(vis))
    return vis


def sobel(rgb, vis=None, gx=None, gy=None):
    rgb = _D(rgb)
    _, H, W = rgb.shape
    gx = np.zeros((H, W)) if gx is None else _D(gx).copy()
    gy = np.zeros((H, W)) if gy is None else _D(gy).copy()
    v = None if vis is None else _U8(vis)
    lib().orc_sobel(W, H, pd(rgb[0]), pd(rgb[1]), pd(rgb[2]), pd(gx), pd(gy), pu8(v))
    return gx, gy


def laplacian(u, vis=None):
    u = _D(u)
    H, W = u.shape
    out = np.zeros((H, W))
    v = None if vis is None else _U8(vis)
    lib().orc_laplacian(W, H, pd(u), pd(out), pu8(v))
    return out


def dtd(u):
    u = _D(u)
    H, W = u.shape
    out = np.zeros((H, W))
    lib().orc_dtd(W, H, pd(u), pd(out))
    return out


def dtg(gx, gy):
    gx, gy = _D(gx), _D(gy)
    H, W = gx.shape
    out = np.zeros((H, W))
    lib().orc_dtg(W, H, pd(gx), pd(gy), pd(out))
    return out


def deform_weights(W, H, tx, ty):
    tx, ty = _F(tx), _F(ty)
    h, w = tx.shape
    out = np.zeros((h, w))
    lib().orc_deform_weights(W, H, w, h, pf(tx), pf(ty), pd(out))
    return out


def angular_weights(gx, gy, sigma, dpx, dpy, thr):
    gx, gy, dpx, dpy = _D(gx), _D(gy), _F(dpx), _F(dpy)
    h, w = gx.shape
    out = np.zeros((h, w))
    lib().orc_angular_weights(w, h, pd(gx), pd(gy), sigma, pf(dpx), pf(dpy), thr, pd(out))
    return out


def decompose_P(P):
    P = _D(P).reshape(12)
    K, R, t, C = np.zeros(9), np.zeros(9), np.zeros(3), np.zeros(3)
    lib().orc_decompose_P(pd(P), pd(K), pd(R), pd(t), pd(C))
    return K.reshape(3, 3), R.reshape(3, 3), t, C


def warps_from_depth(depths, P_views, P_target, dsf, occl_tol=10.0, source_cam=0, sigma_z=10.0):
    depths = _F(depths)
    V, h, w = depths.shape
    P = _D(P_views).reshape(V * 12)
    Pu = _D(P_target).reshape(12)
    tx, ty = np.zeros((V, h, w), np.float32), np.zeros((V, h, w), np.float32)
    dx, dy = np.zeros((V, h, w), np.float32), np.zeros((V, h, w), np.float32)
    ud = np.zeros((h * dsf, w * dsf), np.float32)
    lib().orc_warps_from_depth(V, w, h, dsf, pf(depths), pd(P), pd(Pu), occl_tol, source_cam, sigma_z,
                               pf(tx), pf(ty), pf(dx), pf(dy), pf(ud))
    return tx, ty, dx, dy, ud


def radial_to_ortho(z, K, dzdx=None, dzdy=None):
    """Radial depth (and its x / y derivatives) to orthographic depth
    (fromRadial2Ortho.frag); returns new arrays."""
    z = _F(z).copy()
    H, W = z.shape
    gx = None if dzdx is None else _F(dzdx).copy()
    gy = None if dzdy is None else _F(dzdy).copy()
    Ki = np.linalg.inv(_D(K).reshape(3, 3))
    lib().orc_radial_to_ortho(W, H, pd(_D(Ki).reshape(9)), pf(z), None if gx is None else pf(gx),
                              None if gy is None else pf(gy))
    return z, gx, gy


def gl_target_depth(depths, P_views, P_target, W, H, skip=-1, dzdx=None, dzdy=None):
    """The target depth map of the OpenGL producer (depth splatting, scene.cpp:1506-1544)."""
    depths = _F(depths)
    V, h, w = depths.shape
    gx = None if dzdx is None else _F(dzdx)
    gy = None if dzdy is None else _F(dzdy)
    out = np.zeros((H, W), np.float32)
    lib().orc_gl_target_depth(V, w, h, W, H, pf(depths), None if gx is None else pf(gx), None if gy is None else pf(gy),
                              pd(_D(P_views).reshape(V * 12)), pd(_D(P_target).reshape(12)), skip, pf(out))
    return out


def gl_warps(depths, P_views, P_target, ud, sigma_z=10.0, dzdx=None, dzdy=None):
    """tau, deformation weight and sigma_z dtau/dz of the OpenGL producer
    (tauWarp.frag, tauPartial.frag) against the target depth map ud."""
    depths = _F(depths)
    V, h, w = depths.shape
    H, W = ud.shape
    gx = None if dzdx is None else _F(dzdx)
    gy = None if dzdy is None else _F(dzdy)
    tx, ty, wd, dx, dy = (np.zeros((V, h, w), np.float32) for _ in range(5))
    lib().orc_gl_warps(V, w, h, W, H, pf(depths), None if gx is None else pf(gx), None if gy is None else pf(gy),
                       pd(_D(P_views).reshape(V * 12)), pd(_D(P_target).reshape(12)), pf(_F(ud)), sigma_z,
                       pf(tx), pf(ty), pf(wd), pf(dx), pf(dy))
    return tx, ty, wd, dx, dy


def initial_image(W, H, dsf, tx, ty, views, omega):
    tx, ty, views, omega = _F(tx), _F(ty), _F(views), _D(omega)
    V, h, w = tx.shape
    U = np.zeros((3, H, W))
    nm = np.zeros((H, W))
    lib().orc_initial_image(V, W, H, w, h, dsf, pf(tx), pf(ty), pf(views), pd(omega), pd(U), pd(nm))
    return U, nm


def hole_filling(U, vis, iters=500, eps=1e-3):
    U = _D(U).copy()
    _, H, W = U.shape
    k = lib().orc_hole_filling(W, H, pd(U), pu8(_U8(vis)), iters, eps)
    return U, k


def u_gradient(W, H, dsf, tx, ty, views, omega, sigma):
    tx, ty, views, omega = _F(tx), _F(ty), _F(views), _D(omega)
    V, h, w = tx.shape
    gx, gy, mask, nm = np.zeros((H, W)), np.zeros((H, W)), np.zeros((H, W)), np.zeros((H, W))
    lib().orc_u_gradient(V, W, H, w, h, dsf, pf(tx), pf(ty), pf(views), pd(omega), sigma, pd(gx), pd(gy),
                         pd(mask), pd(nm))
    return gx, gy, mask, nm


def apply_A(W, H, dsf, tx, ty, wdata, alpha, c0, beta, d):
    tx, ty, wdata, d = _F(tx), _F(ty), _D(wdata), _D(d)
    V, h, w = tx.shape
    q = np.zeros((H, W))
    lib().orc_apply_A(V, W, H, w, h, dsf, pf(tx), pf(ty), pd(wdata), alpha, c0, beta, pd(d), pd(q))
    return q


def diag_A(W, H, dsf, tx, ty, wdata, alpha, c0, beta):
    tx, ty, wdata = _F(tx), _F(ty), _D(wdata)
    V, h, w = tx.shape
    dg = np.zeros((H, W))
    lib().orc_diag_A(V, W, H, w, h, dsf, pf(tx), pf(ty), pd(wdata), alpha, c0, beta, pd(dg))
    return dg


def rhs(W, H, dsf, tx, ty, views, wdata, alpha, c0, u0, beta, gx, gy):
    tx, ty, views, wdata = _F(tx), _F(ty), _F(views), _D(wdata)
    V, h, w = tx.shape
    b = np.zeros((3, H, W))
    u0d = None if u0 is None else _D(u0)
    lib().orc_rhs(V, W, H, w, h, dsf, pf(tx), pf(ty), pf(views), pd(wdata), alpha, c0, pd(u0d), beta,
                  pd(_D(gx)), pd(_D(gy)), pd(b))
    return b


def pcg(W, H, dsf, tx, ty, wdata, alpha, c0, beta, B, X0, tol, maxit, precond=True):
    tx, ty, wdata, B = _F(tx), _F(ty), _D(wdata), _D(B)
    V, h, w = tx.shape
    X = _D(X0).copy()
    st = CgStats()
    lib().orc_pcg(V, W, H, w, h, dsf, pf(tx), pf(ty), pd(wdata), alpha, c0, beta, pd(B), pd(X), tol, maxit,
                  int(precond), ctypes.byref(st))
    return X, dict(iterations=list(st.iterations), converged=list(st.converged),
                   rel_residual=list(st.rel_residual), true_rel_residual=list(st.true_rel_residual),
                   status=st.status)


def gibr_energy(W, H, dsf, tx, ty, views, wdata, alpha, c0, u0, beta, gx, gy, U):
    tx, ty, views, wdata = _F(tx), _F(ty), _F(views), _D(wdata)
    V, h, w = tx.shape
    u0d = None if u0 is None else _D(u0)
    return lib().orc_gibr_energy(V, W, H, w, h, dsf, pf(tx), pf(ty), pf(views), pd(wdata), alpha, c0, pd(u0d),
                                 beta, pd(_D(gx)), pd(_D(gy)), pd(_D(U)))


def pd_dual(U, PX, PY, sigma, g=None):
    U = _D(U)
    PX, PY = _D(PX).copy(), _D(PY).copy()
    _, H, W = U.shape
    gg = None if g is None else _D(g)
    lib().orc_pd_dual(W, H, sigma, pd(U), pd(PX), pd(PY), pd(gg))
    return PX, PY


def pd_primal(u, f, px, py, tau, lam):
    u = _D(u).copy()
    H, W = u.shape
    lib().orc_pd_primal(W, H, tau, lam, pd(u), pd(_D(f)), pd(_D(px)), pd(_D(py)))
    return u


def gauss_smooth(img, taps=11, sigma=2.0):
    img = _D(img)
    H, W = img.shape
    out = np.zeros((H, W))
    lib().orc_gauss_smooth(W, H, taps, sigma, pd(img), pd(out))
    return out


def regularizer_weight(W, H, dsf, tx, ty, omega, sigma, lmax, lmin, norm_mask):
    tx, ty, omega = _F(tx), _F(ty), _D(omega)
    V, h, w = tx.shape
    nm = _D(norm_mask).copy()
    g = np.zeros((H, W))
    lib().orc_regularizer_weight(V, W, H, w, h, dsf, pf(tx), pf(ty), pd(omega), sigma, lmax, lmin, pd(nm), pd(g))
    return g, nm


class SrState:
    """Mutable state of the IBR_color solver in FP64 (mirrors coco_vtv_workspace)."""

    def __init__(self, W, H, U):
        self.U = _D(U).copy()
        self.Uq = self.U.copy()
        self.PX = np.zeros((3, H, W))
        self.PY = np.zeros((3, H, W))
        self.norm_mask = np.zeros((H, W))
        self.fista_alpha = ctypes.c_double(1.0)
        self.G = np.zeros((3, H, W))


def fista_step(W, H, dsf, tx, ty, views, omega, sigma, dt_alpha, dt_beta, gradient_step, gw_type, lam, inner,
               g, state):
    tx, ty, views, omega = _F(tx), _F(ty), _F(views), _D(omega)
    V, h, w = tx.shape
    gg = None if g is None else _D(g)
    lib().orc_fista_step(V, W, H, w, h, dsf, pf(tx), pf(ty), pf(views), pd(omega), sigma, dt_alpha, dt_beta,
                         gradient_step, gw_type, lam, inner, pd(gg), pd(state.U), pd(state.Uq), pd(state.PX),
                         pd(state.PY), pd(state.norm_mask), ctypes.byref(state.fista_alpha), pd(state.G))


def sr_energy(W, H, dsf, tx, ty, views, omega, sigma, lam, dt_alpha, dt_beta, gw_type, vis, U):
    tx, ty, views, omega = _F(tx), _F(ty), _F(views), _D(omega)
    V, h, w = tx.shape
    parts = np.zeros(3)
    e = lib().orc_sr_energy(V, W, H, w, h, dsf, pf(tx), pf(ty), pf(views), pd(omega), sigma, lam, dt_alpha,
                            dt_beta, gw_type, pu8(_U8(vis)), pd(_D(U)), pd(parts))
    return e, parts


def push_pull(U, vis, levels=10):
    U = _D(U).copy()
    _, H, W = U.shape
    lib().orc_push_pull(W, H, levels, pd(U), pu8(_U8(vis)))
    return U


def psnr(A, B):
    A, B = _D(A), _D(B)
    _, H, W = A.shape
    return lib().orc_psnr(W, H, pd(A), pd(B))


# --------------------------------------------------------------------------
# Pipelines composed from the functions above, in the order of the GPU path
# --------------------------------------------------------------------------

def gibr_weights(W, H, dsf, tx, ty, dpx, dpy, U, sigma, thr):
    """compute_weights (gradientIBR.cu:558-648) for every view, FP64."""
    gxs, gys = sobel(U)
    V = tx.shape[0]
    om = np.zeros(tx.shape)
    for v in range(V):
        g = gather(W, H, dsf, tx[v], ty[v], np.stack([gxs, gys]))
        ang = angular_weights(g[0], g[1], sigma, dpx[v], dpy[v], thr)
        om[v] = ang * deform_weights(W, H, tx[v], ty[v])
    return om


def gibr_pipeline(views, depths, P_views, P_target, dsf, sigma, thr, alpha, beta, tol, maxit,
                  occl_tol=10.0, source_cam=0, sigma_z=10.0, precond=True, taus=None):
    """The gradient-constraint IBR solve end to end in FP64: warps from depth,
    visibility, weights at U = 0, weighted initial image + Poisson hole filling,
    target gradient, PCG of the normal equations.  `taus` = (tx, ty, dpx, dpy)
    skips the warp stage (stage-wise parity with the GPU's warps)."""
    V, _, h, w = views.shape
    W, H = w * dsf, h * dsf
    if taus is None:
        tx, ty, dpx, dpy, _ = warps_from_depth(depths, P_views, P_target, dsf, occl_tol, source_cam, sigma_z)
    else:
        tx, ty, dpx, dpy = taus
    vis = target_visibility(W, H, dsf, list(zip(tx, ty)))
    omega = gibr_weights(W, H, dsf, tx, ty, dpx, dpy, np.zeros((3, H, W)), sigma, thr)
    U0, _ = initial_image(W, H, dsf, tx, ty, views, omega)
    U0, hole_iters = hole_filling(U0, vis)
    gx, gy, mask, _ = u_gradient(W, H, dsf, tx, ty, views, omega, sigma)
    wdata = omega * sigma * sigma
    b = rhs(W, H, dsf, tx, ty, views, wdata, alpha, 0.0, None, beta, gx, gy)
    X, st = pcg(W, H, dsf, tx, ty, wdata, alpha, 0.0, beta, b, U0, tol, maxit, precond)
    E = gibr_energy(W, H, dsf, tx, ty, views, wdata, alpha, 0.0, None, beta, gx, gy, X)
    return dict(U=X, U0=U0, stats=st, energy=E, tx=tx, ty=ty, dpx=dpx, dpy=dpy, omega=omega, vis=vis, gx=gx,
                gy=gy, mask=mask, b=b, hole_iters=hole_iters)


def sr_weights(W, H, dsf, tx, ty, dpx, dpy, U, vis, PX0, PY0, sigma, thr):
    """coco_vtv_sr_compute_weights_unstructured (optimization.cu:1066-1167): the
    gated Sobel of U is written into the channel-0 dual variables (the
    reference's scratch choice, :1090-1091) and gathered from there.
    Returns omega and the updated (PX0, PY0)."""
    gx, gy = sobel(U, vis, PX0, PY0)
    V = tx.shape[0]
    om = np.zeros(tx.shape)
    for v in range(V):
        g = gather(W, H, dsf, tx[v], ty[v], np.stack([gx, gy]))
        ang = angular_weights(g[0], g[1], sigma, dpx[v], dpy[v], thr)
        om[v] = ang * deform_weights(W, H, tx[v], ty[v])
    return om, gx, gy


def coverage(W, H, dsf, tx, ty):
    cov = np.zeros((H, W))
    for v in range(tx.shape[0]):
        splat(W, H, dsf, tx[v], ty[v], None, wsum=cov)
    return cov


# --------------------------------------------------------------------------
# The warm-start U-Net (lfrv_oracle_cnn.c)
# --------------------------------------------------------------------------

def conv3x3(X, Wt, b=None):
    """Y[n, y, x, co] = sum_{t, ci} Wt[co, t, ci] X[n, y+dy, x+dx, ci] (+ b), zero padding;
    X (B, H, W, Cin), Wt (Cout, 9, Cin)."""
    X, Wt = _D(X), _D(Wt)
    B, H, W, Cin = X.shape
    Cout = Wt.shape[0]
    Y = np.zeros((B, H, W, Cout))
    lib().orc_conv3x3(B, H, W, Cin, Cout, pd(X), pd(Wt), pd(None if b is None else _D(b)), pd(Y))
    return Y


def cnn_num_params(levels, base, cpad):
    return int(lib().orc_cnn_num_params(levels, base, cpad))


def cnn_forward_backward(levels, base_ch, x, params, base, target, round_bf16=False, grads=True):
    """x (B, H, W, cpad); base, target (B, 3, H, W).  Returns (loss, pred, grads or None)."""
    x = _D(x)
    B, H, W, cpad = x.shape
    params = _D(params)
    pred = np.zeros((B, 3, H, W))
    g = np.zeros(params.size) if grads else None
    loss = lib().orc_cnn_forward_backward(levels, base_ch, cpad, B, H, W, pd(x), pd(params), pd(_D(base)),
                                          pd(_D(target)), int(round_bf16), pd(pred), pd(g))
    return loss, pred, g
