"""The device PSNR (lfrv_op_psnr) against the host one (gsl_image_psnr as
lfrv_psnr computes it) and the oracle's."""
import ctypes

import numpy as np
import pytest

import lfrv
import oracle

pytestmark = pytest.mark.gpu


def op_psnr(a, b, fn="lfrv_op_psnr"):
    """PSNR of two planar RGB images through the device operator `fn`."""
    L = lfrv.lib()
    f = getattr(L, fn)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    H, W = a.shape[1], a.shape[2]
    da, db = lfrv.DeviceArray(a), lfrv.DeviceArray(b)
    dout = lfrv.DeviceArray((2,))  # one double
    st = f(W, H, da.ptr, db.ptr, dout.ptr, None)
    lfrv.device_sync()
    out = np.zeros(1, np.float64)
    assert L.lfrv_memcpy(out.ctypes.data, dout.ptr, 8) == 0
    for d in (da, db, dout):
        d.free()
    return st, float(out[0])


@pytest.mark.parametrize("shape", [(5, 5), (37, 29), (256, 320), (1031, 1025)])
def test_device_psnr_matches_host(shape):
    H, W = shape
    print(f"lfrv_op_psnr {W}x{H}", flush=True)
    rng = np.random.default_rng(H * 7 + W)
    a = rng.random((3, H, W)).astype(np.float32)
    b = np.clip(a + 0.03 * rng.standard_normal(a.shape), 0, 1).astype(np.float32)
    st, p = op_psnr(a, b)
    assert st == lfrv.OK
    assert abs(p - lfrv.psnr(a, b)) < 1e-9 * abs(p)
    assert abs(p - oracle.psnr(a, b)) < 1e-9 * abs(p)


def test_device_psnr_of_equal_images_and_bad_sizes():
    a = np.random.default_rng(1).random((3, 12, 14)).astype(np.float32)
    b = a.copy()
    b[:, :2, :] = 0.0  # differences only in the 2-px border
    st, p = op_psnr(a, b)
    assert st == lfrv.OK and p == np.finfo(np.float64).max
    st, _ = op_psnr(a[:, :4, :].copy(), a[:, :4, :].copy())
    assert st == lfrv.E_ARG
