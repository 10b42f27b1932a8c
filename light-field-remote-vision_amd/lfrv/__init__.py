"""ctypes bindings of liblfrv.so, the MI355X-native variational light-field solver.

The C-ABI is declared in include/lfrv.h and include/lfrv_synth.h.  This module
mirrors it one-to-one with numpy arrays in and out (host memory; the library
copies to and from the device itself).  Nothing here computes: every numeric
call goes to the HIP library, and a missing library raises immediately.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "liblfrv.so")

MODE_GRADIENT_IBR = 0
MODE_COLOR_SR = 1

OK, E_ARG, E_HIP, E_NONCONVERGED, E_NAN, E_STATE, E_BREAKDOWN = range(7)
STATUS_NAMES = {0: "OK", 1: "E_ARG", 2: "E_HIP", 3: "E_NONCONVERGED", 4: "E_NAN", 5: "E_STATE", 6: "E_BREAKDOWN"}

BUF = dict(
    TAU_X=0, TAU_Y=1, DPART_X=2, DPART_Y=3, OMEGA=4, U=5, UQ=6, G=7, PX=8, PY=9, REG_G=10, NORM_MASK=11,
    COVERAGE=12, U_GRAD_X=13, U_GRAD_Y=14, TARGET_MASK=15, VISIBILITY=16, TARGET_DEPTH=17, CG_B=18, CG_MINV=19,
    TAU_WEIGHT=20,
)

TIMING_CLASSES = dict(cg_op=0, cg_update=1, pd_step=2, dataterm=3, warp=4, splat=5)


class LfrvError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"lfrv {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int), ("nviews", ctypes.c_int), ("w", ctypes.c_int), ("h", ctypes.c_int),
        ("dsf", ctypes.c_int), ("sigma_sensor", ctypes.c_float), ("ugrad_threshold", ctypes.c_float),
        ("dw_type", ctypes.c_int), ("gw_type", ctypes.c_int), ("dt_alpha", ctypes.c_float),
        ("dt_beta", ctypes.c_float), ("gradient_step", ctypes.c_float), ("lambda_", ctypes.c_float),
        ("lambda_max_factor", ctypes.c_float), ("inner_iterations", ctypes.c_int),
        ("occlusion_tol_px", ctypes.c_float), ("sigma_z", ctypes.c_float),
        ("backproject_source_cam", ctypes.c_int), ("tile_w", ctypes.c_int), ("tile_h", ctypes.c_int),
        ("pull_lds_kb", ctypes.c_int),
    ]


class CgStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int * 3), ("converged", ctypes.c_int * 3),
        ("rel_residual", ctypes.c_double * 3), ("true_rel_residual", ctypes.c_double * 3),
        ("status", ctypes.c_int), ("ms", ctypes.c_double),
    ]

    def as_dict(self):
        return dict(iterations=list(self.iterations), converged=list(self.converged),
                    rel_residual=list(self.rel_residual), true_rel_residual=list(self.true_rel_residual),
                    status=self.status, ms=self.ms)


def _with_status(stats, status):
    """The stats as a dict whose status is the call's own: a call rejected
    before the solve (bad arguments, wrong state) leaves the struct unfilled."""
    d = stats.as_dict()
    d["status"] = status
    return d


class SrStats(ctypes.Structure):
    _fields_ = [
        ("meta_iterations", ctypes.c_int), ("fista_steps", ctypes.c_int), ("last_stop", ctypes.c_int),
        ("converged", ctypes.c_int), ("energy", ctypes.c_float), ("ms", ctypes.c_double),
    ]

    def as_dict(self):
        return dict(meta_iterations=self.meta_iterations, fista_steps=self.fista_steps, last_stop=self.last_stop,
                    converged=self.converged, energy=self.energy, ms=self.ms)


class SynthDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("seed", ctypes.c_uint64), ("grid", ctypes.c_int), ("w", ctypes.c_int),
        ("h", ctypes.c_int), ("dsf", ctypes.c_int), ("render_scale", ctypes.c_int), ("focal", ctypes.c_double),
        ("baseline", ctypes.c_double), ("pp_shift", ctypes.c_double), ("plane_depth", ctypes.c_double),
        ("nplanes", ctypes.c_int), ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("random_target", ctypes.c_int), ("noise_sigma", ctypes.c_double), ("tex_min_period", ctypes.c_double),
        ("tex_max_period", ctypes.c_double),
    ]


_lib = None
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def lib():
    """Load liblfrv.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"liblfrv.so not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    st = ctypes.c_int
    sig = {
        "lfrv_default_params": (None, [ctypes.POINTER(Params), st, st, st, st, st]),
        "lfrv_create": (st, [ctypes.POINTER(Params), ctypes.POINTER(_vp)]),
        "lfrv_destroy": (None, [_vp]),
        "lfrv_last_error": (ctypes.c_char_p, [_vp]),
        "lfrv_sync": (st, [_vp]),
        "lfrv_set_view": (st, [_vp, st, _vp]),
        "lfrv_set_tau": (st, [_vp, st, _vp, _vp]),
        "lfrv_set_partial_tau": (st, [_vp, st, _vp, _vp, _vp]),
        "lfrv_set_depth": (st, [_vp, st, _vp, _dp]),
        "lfrv_set_target_camera": (st, [_vp, _dp]),
        "lfrv_warps_from_depth": (st, [_vp]),
        "lfrv_set_depth_gradient": (st, [_vp, st, _vp, _vp]),
        "lfrv_warps_from_target_depth": (st, [_vp, st]),
        "lfrv_prepare": (st, [_vp]),
        "lfrv_setup_visibility": (st, [_vp]),
        "lfrv_compute_weights": (st, [_vp]),
        "lfrv_compute_initial_image": (st, [_vp, _vp]),
        "lfrv_get_solution": (st, [_vp, _vp]),
        "lfrv_get_hole_stats": (st, [_vp, ctypes.POINTER(st), ctypes.POINTER(st)]),
        "lfrv_set_solution": (st, [_vp, _vp]),
        "lfrv_clear_solution": (st, [_vp]),
        "lfrv_init_u_gradient": (st, [_vp]),
        "lfrv_cg_solve": (st, [_vp, ctypes.c_double, st, ctypes.POINTER(CgStats)]),
        "lfrv_ibr_gradient": (st, [_vp, ctypes.c_double, st, ctypes.POINTER(CgStats)]),
        "lfrv_poisson_cg": (st, [_vp, ctypes.c_double, ctypes.c_double, st, st, ctypes.POINTER(CgStats)]),
        "lfrv_gibr_energy": (st, [_vp, _dp]),
        "lfrv_sr_init_regularizer_weight": (st, [_vp]),
        "lfrv_sr_iteration_fista": (st, [_vp]),
        "lfrv_sr_iterate": (st, [_vp, st]),
        "lfrv_sr_primal_energy": (st, [_vp, _dp, _dp]),
        "lfrv_sr_count_moving": (st, [_vp, _vp, ctypes.c_float, ctypes.POINTER(ctypes.c_int64)]),
        "lfrv_sr_solve": (st, [_vp, st, st, ctypes.POINTER(SrStats)]),
        "lfrv_get_buffer": (st, [_vp, st, _vp, ctypes.POINTER(ctypes.c_size_t)]),
        "lfrv_set_buffer": (st, [_vp, st, _vp]),
        "lfrv_set_timing": (st, [_vp, st]),
        "lfrv_get_timing": (st, [_vp, st, _dp, ctypes.POINTER(ctypes.c_int64)]),
        "lfrv_get_bin_stats": (st, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]),
        "lfrv_op_gather": (st, [st, st, st, st, st, _vp, _vp, _vp, st, _vp, _vp]),
        "lfrv_op_splat_atomic": (st, [st, st, st, st, st, _vp, _vp, _vp, st, _vp, _vp, _vp, _vp]),
        "lfrv_op_sobel": (st, [st, st, _vp, _vp, _vp, _vp, _vp]),
        "lfrv_op_radial_to_ortho": (st, [st, st, _dp, _vp, _vp, _vp, _vp]),
        "lfrv_psnr": (ctypes.c_double, [st, st, _vp, _vp]),
        "lfrv_dev_alloc": (st, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
        "lfrv_dev_free": (st, [_vp]),
        "lfrv_dev_copy_kernel": (st, [_vp, _vp, ctypes.c_size_t, _vp]),
        "lfrv_memcpy": (st, [_vp, _vp, ctypes.c_size_t]),
        "lfrv_device_sync": (st, []),
        "lfrv_synth_preset": (None, [ctypes.POINTER(SynthDesc), st, ctypes.c_uint64]),
        "lfrv_synth_nviews": (st, [ctypes.POINTER(SynthDesc)]),
        "lfrv_synth_generate": (st, [ctypes.POINTER(SynthDesc), _vp, _vp, _dp, _dp, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    """Address of an array the caller keeps referenced for the whole C call."""
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _in(a):
    """(array, address) of `a` as contiguous float32.  When `a` needs a
    conversion the array is a new temporary: the caller holds the returned
    pair until the C call has returned, so the address stays valid (a bare
    address of a temporary would point at freed memory during the call)."""
    if a is None:
        return None, None
    arr = _f32(a)
    return arr, ctypes.c_void_p(arr.ctypes.data)


def default_params(mode, nviews, w, h, dsf, **kw):
    p = Params()
    lib().lfrv_default_params(ctypes.byref(p), mode, nviews, w, h, dsf)
    for k, v in kw.items():
        setattr(p, "lambda_" if k == "lambda" else k, v)
    return p


class Context:
    """One novel-view problem on the GPU (an lfrv_ctx)."""

    def __init__(self, params):
        self.params = params
        self._h = ctypes.c_void_p()
        st = lib().lfrv_create(ctypes.byref(params), ctypes.byref(self._h))
        if st != OK:
            raise LfrvError(st, "lfrv_create failed")
        self.V, self.w, self.h, self.dsf = params.nviews, params.w, params.h, params.dsf
        self.W, self.H = self.w * self.dsf, self.h * self.dsf

    def close(self):
        if self._h:
            lib().lfrv_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        if st != OK:
            raise LfrvError(st, lib().lfrv_last_error(self._h).decode())

    # inputs
    def set_view(self, i, rgb):
        a, pa = _in(rgb)
        self._check(lib().lfrv_set_view(self._h, i, pa))

    def set_tau(self, i, tx, ty):
        a, pa = _in(tx)
        b, pb = _in(ty)
        self._check(lib().lfrv_set_tau(self._h, i, pa, pb))

    def set_partial_tau(self, i, sigma_z, dx, dy):
        a, pa = _in(sigma_z)
        b, pb = _in(dx)
        c, pc = _in(dy)
        self._check(lib().lfrv_set_partial_tau(self._h, i, pa, pb, pc))

    def set_depth(self, i, depth, P):
        Pd = np.ascontiguousarray(P, dtype=np.float64).reshape(12)
        a, pa = _in(depth)
        self._check(lib().lfrv_set_depth(self._h, i, pa, Pd.ctypes.data_as(_dp)))

    def set_target_camera(self, P):
        Pd = np.ascontiguousarray(P, dtype=np.float64).reshape(12)
        self._check(lib().lfrv_set_target_camera(self._h, Pd.ctypes.data_as(_dp)))

    # steps
    def warps_from_depth(self):
        self._check(lib().lfrv_warps_from_depth(self._h))

    def set_depth_gradient(self, i, dzdx=None, dzdy=None):
        """dz/dx, dz/dy of view i for the OpenGL warp producer (None: zero)."""
        if dzdx is None:
            self._check(lib().lfrv_set_depth_gradient(self._h, i, None, None))
        else:
            a, pa = _in(dzdx)
            b, pb = _in(dzdy)
            self._check(lib().lfrv_set_depth_gradient(self._h, i, pa, pb))

    def warps_from_target_depth(self, skip_view=-1):
        """Target depth by depth splatting, then tau / dpart as the OpenGL exporter computes them."""
        self._check(lib().lfrv_warps_from_target_depth(self._h, skip_view))

    def prepare(self):
        self._check(lib().lfrv_prepare(self._h))

    def setup_visibility(self):
        self._check(lib().lfrv_setup_visibility(self._h))

    def compute_weights(self):
        self._check(lib().lfrv_compute_weights(self._h))

    def compute_initial_image(self, u0=None):
        a, pa = _in(u0)
        self._check(lib().lfrv_compute_initial_image(self._h, pa))

    def init_u_gradient(self):
        self._check(lib().lfrv_init_u_gradient(self._h))

    def cg_solve(self, tol=1e-5, maxit=1000, check=True):
        s = CgStats()
        st = lib().lfrv_cg_solve(self._h, tol, maxit, ctypes.byref(s))
        if check:
            self._check(st)
        return _with_status(s, st)

    def ibr_gradient(self, tol=1e-5, maxit=5000, check=True):
        """The whole device-resident IBR_gradient sequence in one call (lfrv_ibr_gradient)."""
        s = CgStats()
        st = lib().lfrv_ibr_gradient(self._h, tol, maxit, ctypes.byref(s))
        if check:
            self._check(st)
        return _with_status(s, st)

    def poisson_cg(self, lam=1.0, tol=1e-5, maxit=1000, precond=False, check=True):
        s = CgStats()
        st = lib().lfrv_poisson_cg(self._h, lam, tol, maxit, int(precond), ctypes.byref(s))
        if check:
            self._check(st)
        return _with_status(s, st)

    def gibr_energy(self):
        e = ctypes.c_double()
        self._check(lib().lfrv_gibr_energy(self._h, ctypes.byref(e)))
        return e.value

    def sr_init_regularizer_weight(self):
        self._check(lib().lfrv_sr_init_regularizer_weight(self._h))

    def sr_iteration_fista(self):
        self._check(lib().lfrv_sr_iteration_fista(self._h))

    def sr_primal_energy(self):
        e = ctypes.c_double()
        parts = (ctypes.c_double * 3)()
        self._check(lib().lfrv_sr_primal_energy(self._h, ctypes.byref(e), parts))
        return e.value, list(parts)

    def sr_count_moving(self, prev, max_diff=1e-3):
        m = ctypes.c_int64()
        a, pa = _in(prev)
        self._check(lib().lfrv_sr_count_moving(self._h, pa, max_diff, ctypes.byref(m)))
        return m.value

    def sr_iterate(self, steps):
        """`steps` FISTA steps queued in one call (lfrv_sr_iterate)."""
        self._check(lib().lfrv_sr_iterate(self._h, steps))

    def sr_solve(self, niter, nmeta):
        """The re-weighted FISTA loop with the driver's stopping rules on the device (lfrv_sr_solve)."""
        s = SrStats()
        self._check(lib().lfrv_sr_solve(self._h, niter, nmeta, ctypes.byref(s)))
        return s.as_dict()

    def hole_stats(self):
        """(hole pixels, diffusion sweeps) of the last Poisson hole filling."""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._check(lib().lfrv_get_hole_stats(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def sync(self):
        self._check(lib().lfrv_sync(self._h))

    # outputs
    def get_solution(self):
        out = np.empty(3 * self.W * self.H, dtype=np.float32)
        self._check(lib().lfrv_get_solution(self._h, _ptr(out)))
        return out.reshape(3, self.H, self.W)

    def set_solution(self, rgb):
        a, pa = _in(rgb)
        self._check(lib().lfrv_set_solution(self._h, pa))

    def clear_solution(self):
        self._check(lib().lfrv_clear_solution(self._h))

    def get_buffer(self, name):
        which = BUF[name] if isinstance(name, str) else name
        cnt = ctypes.c_size_t()
        self._check(lib().lfrv_get_buffer(self._h, which, None, ctypes.byref(cnt)))
        out = np.empty(cnt.value, dtype=np.float32)
        self._check(lib().lfrv_get_buffer(self._h, which, _ptr(out), ctypes.byref(cnt)))
        return out

    def set_buffer(self, name, arr):
        which = BUF[name] if isinstance(name, str) else name
        a, pa = _in(arr)
        self._check(lib().lfrv_set_buffer(self._h, which, pa))

    def set_timing(self, enable=True):
        self._check(lib().lfrv_set_timing(self._h, int(enable)))

    def get_timing(self, kclass):
        k = TIMING_CLASSES[kclass] if isinstance(kclass, str) else kclass
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(lib().lfrv_get_timing(self._h, k, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def bin_stats(self):
        e = ctypes.c_int64()
        t = ctypes.c_int()
        self._check(lib().lfrv_get_bin_stats(self._h, ctypes.byref(e), ctypes.byref(t)))
        return e.value, t.value


class DeviceArray:
    """A float32 device buffer owned by the library's HIP runtime."""

    def __init__(self, shape_or_array):
        if isinstance(shape_or_array, np.ndarray):
            self.shape = shape_or_array.shape
        else:
            self.shape = tuple(shape_or_array)
        self.nbytes = int(np.prod(self.shape)) * 4
        self.ptr = ctypes.c_void_p()
        if lib().lfrv_dev_alloc(self.nbytes, ctypes.byref(self.ptr)) != OK:
            raise LfrvError(E_HIP, "lfrv_dev_alloc failed")
        if isinstance(shape_or_array, np.ndarray):
            self.upload(shape_or_array)

    def upload(self, a):
        a = _f32(a)
        assert a.nbytes == self.nbytes
        if lib().lfrv_memcpy(self.ptr, _ptr(a), self.nbytes) != OK:
            raise LfrvError(E_HIP, "lfrv_memcpy failed")

    def numpy(self):
        out = np.empty(self.shape, np.float32)
        if lib().lfrv_memcpy(_ptr(out), self.ptr, self.nbytes) != OK:
            raise LfrvError(E_HIP, "lfrv_memcpy failed")
        return out

    def free(self):
        if self.ptr:
            lib().lfrv_dev_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_sync():
    if lib().lfrv_device_sync() != OK:
        raise LfrvError(E_HIP, "lfrv_device_sync failed")


def copy_kernel(dst, src):
    """Device-to-device copy of src into dst (DeviceArrays of equal size) by
    the library's plain streaming copy kernel, on the default stream."""
    if dst.nbytes != src.nbytes:
        raise LfrvError(E_ARG, "copy_kernel: size mismatch")
    if lib().lfrv_dev_copy_kernel(dst.ptr, src.ptr, src.nbytes, None) != OK:
        raise LfrvError(E_HIP, "lfrv_dev_copy_kernel failed")


def psnr(a, b):
    a = _f32(a)
    b = _f32(b)
    H, W = a.shape[-2], a.shape[-1]
    return lib().lfrv_psnr(W, H, _ptr(a), _ptr(b))


def synth_desc(config, seed, **overrides):
    d = SynthDesc()
    lib().lfrv_synth_preset(ctypes.byref(d), config, seed)
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


def synth(desc_or_config, seed=0, **overrides):
    """Render a synthetic light field; returns a dict of numpy arrays."""
    d = desc_or_config if isinstance(desc_or_config, SynthDesc) else synth_desc(desc_or_config, seed, **overrides)
    V = lib().lfrv_synth_nviews(ctypes.byref(d))
    w, h, W, H = d.w, d.h, d.w * d.dsf, d.h * d.dsf
    views = np.zeros((V, 3, h, w), np.float32)
    depths = np.zeros((V, h, w), np.float32)
    Pv = np.zeros((V, 3, 4), np.float64)
    Pu = np.zeros((3, 4), np.float64)
    tgt = np.zeros((3, H, W), np.float32)
    tdep = np.zeros((H, W), np.float32)
    st = lib().lfrv_synth_generate(ctypes.byref(d), _ptr(views), _ptr(depths), Pv.ctypes.data_as(_dp),
                                   Pu.ctypes.data_as(_dp), _ptr(tgt), _ptr(tdep))
    if st != 0:
        raise LfrvError(E_ARG, "lfrv_synth_generate failed")
    return dict(desc=d, views=views, depths=depths, P_views=Pv, P_target=Pu, target=tgt, target_depth=tdep,
                V=V, w=w, h=h, W=W, H=H, dsf=d.dsf)


def radial_to_ortho(z, K, dzdx=None, dzdy=None):
    """The OpenGL path's depth-map filter (lfrv_op_radial_to_ortho) on the GPU:
    radial depth and its x / y derivatives to orthographic depth.  Returns
    numpy arrays (the derivatives None where not given)."""
    K = np.ascontiguousarray(K, np.float64).reshape(9)
    H, W = np.asarray(z).shape
    bufs = [DeviceArray(_f32(a)) if a is not None else None for a in (z, dzdx, dzdy)]
    st = lib().lfrv_op_radial_to_ortho(W, H, K.ctypes.data_as(_dp), bufs[0].ptr, bufs[1].ptr if bufs[1] else None,
                                       bufs[2].ptr if bufs[2] else None, None)
    if st != OK:
        raise LfrvError(st, "lfrv_op_radial_to_ortho failed")
    device_sync()
    out = [b.numpy() if b is not None else None for b in bufs]
    for b in bufs:
        if b is not None:
            b.free()
    return tuple(out)
