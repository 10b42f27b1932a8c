"""lfrv_op_psnr_noalloc (the device PSNR with its partial sums in a fixed
device array) against lfrv_op_psnr (the same kernels over memory it
allocates), on two calls in a row, the second with a smaller image, and its
refusal of an image beyond the fixed array."""
import numpy as np
import pytest

import lfrv
from test_gpu_quality import op_psnr

pytestmark = pytest.mark.gpu


def test_noalloc_psnr_equals_the_allocating_one():
    rng = np.random.default_rng(11)
    for H, W in ((700, 901), (33, 21)):        # the second call smaller: stale partials must not leak in
        print(f"lfrv_op_psnr_noalloc {W}x{H}", flush=True)
        a = rng.random((3, H, W)).astype(np.float32)
        b = np.clip(a + 0.02 * rng.standard_normal(a.shape), 0, 1).astype(np.float32)
        st0, p0 = op_psnr(a, b, "lfrv_op_psnr")
        st1, p1 = op_psnr(a, b, "lfrv_op_psnr_noalloc")
        assert st0 == lfrv.OK and st1 == lfrv.OK
        assert p1 == p0                           # same kernels, same order of the sums: bit-identical
        assert abs(p1 - lfrv.psnr(a, b)) < 1e-9 * p1


def test_noalloc_psnr_refuses_images_beyond_its_scratch():
    # 16384 partials of 4096 interior pixels: an image of 8196 x 8197 has
    # 8192 x 8193 interior pixels, one block more than the array holds; the
    # size is refused before any device pointer is read
    import ctypes
    L = lfrv.lib()
    f = L.lfrv_op_psnr_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    assert f(8196, 8197, ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16), None) == lfrv.E_ARG
    assert f(4, 100, ctypes.c_void_p(16), ctypes.c_void_p(16), ctypes.c_void_p(16), None) == lfrv.E_ARG
