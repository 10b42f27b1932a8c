"""ctypes bindings of liblfrv_host.so: the reference's file formats and configuration.

Declared in include/lfrv_io.h.  PFM and 8-bit PNG images as planar float RGB
(3, h, w); the coco::config / Config_data parse of a reference command line.
The command-line driver built beside the library is lfrv_cli (CLI_PATH).
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB_PATH = os.path.join(_PKG, "liblfrv_host.so")
CLI_PATH = os.path.join(_PKG, "lfrv_cli")

_lib = None


def lib():
    """Load liblfrv_host.so (built in-tree with liblfrv.so)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HOST_LIB_PATH):
        raise RuntimeError(f"liblfrv_host.so not built: {HOST_LIB_PATH} is missing (run __graft_entry__.build())")
    L = ctypes.CDLL(HOST_LIB_PATH)
    sz = ctypes.POINTER(ctypes.c_size_t)
    L.lfrv_io_load_image.restype = ctypes.c_int
    L.lfrv_io_load_image.argtypes = [ctypes.c_char_p, sz, sz, ctypes.c_void_p]
    L.lfrv_io_save_image.restype = ctypes.c_int
    L.lfrv_io_save_image.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p]
    L.lfrv_io_config_data.restype = ctypes.c_int
    L.lfrv_io_config_data.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_char_p, ctypes.c_size_t]
    L.lfrv_io_export_warps.restype = ctypes.c_int
    L.lfrv_io_export_warps.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    L.lfrv_io_from_mve.restype = ctypes.c_int
    L.lfrv_io_from_mve.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
    _dp = ctypes.POINTER(ctypes.c_double)
    _fp = ctypes.POINTER(ctypes.c_float)
    L.lfrv_io_mve_camera.restype = ctypes.c_int
    L.lfrv_io_mve_camera.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
    L.lfrv_io_mve_depth.restype = ctypes.c_int
    L.lfrv_io_mve_depth.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _fp, _fp, _fp]
    _lib = L
    return L


def load_image(path):
    """(3, h, w) float32 planes of a .pfm or .png file."""
    w, h = ctypes.c_size_t(), ctypes.c_size_t()
    if lib().lfrv_io_load_image(os.fsencode(path), ctypes.byref(w), ctypes.byref(h), None) != 0:
        raise IOError(f"cannot read {path}")
    out = np.zeros((3, h.value, w.value), np.float32)
    if lib().lfrv_io_load_image(os.fsencode(path), ctypes.byref(w), ctypes.byref(h),
                                ctypes.c_void_p(out.ctypes.data)) != 0:
        raise IOError(f"cannot read {path}")
    return out


def save_image(path, rgb):
    """Write (3, h, w) planes as .pfm or .png (by extension)."""
    a = np.ascontiguousarray(rgb, np.float32)
    if a.ndim != 3 or a.shape[0] != 3:
        raise ValueError("expected (3, h, w) planes")
    if lib().lfrv_io_save_image(os.fsencode(path), a.shape[2], a.shape[1], ctypes.c_void_p(a.ctypes.data)) != 0:
        raise IOError(f"cannot write {path}")


def config_data(argv):
    """Config_data fields (as strings) parsed from a reference command line
    (argv[0] is the program name)."""
    args = (ctypes.c_char_p * len(argv))(*[os.fsencode(a) for a in argv])
    n = lib().lfrv_io_config_data(len(argv), args, None, 0)
    if n < 0:
        raise ValueError("bad command line")
    buf = ctypes.create_string_buffer(n)
    lib().lfrv_io_config_data(len(argv), args, buf, n)
    out = {}
    for line in buf.value.decode().splitlines():
        k, _, v = line.partition(" ")
        out[k] = v
    return out


def export_warps(argv):
    """lfrv_ref::export_warps for a reference command line (argv[0] the program
    name): the OpenGL path's target depth map and warps, computed on the GPU
    and written as the tau / dpart / uDepth PFMs the config names."""
    args = (ctypes.c_char_p * len(argv))(*[os.fsencode(a) for a in argv])
    if lib().lfrv_io_export_warps(len(argv), args) != 0:
        raise RuntimeError("export_warps failed (reason on stderr)")


def from_mve(argv):
    """lfrv_ref::from_mve_to_coco (the fromMVE2coco algorithm) for a reference
    command line (argv[0] the program name)."""
    args = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
    if lib().lfrv_io_from_mve(len(argv), args) != 0:
        raise RuntimeError("from_mve failed (reason on stderr)")


def mve_camera(meta_ini, w, h):
    """(K, R, t, P) of an MVE meta.ini's [camera] for views of w x h, or None."""
    K, R, t, P = np.zeros(9), np.zeros(9), np.zeros(3), np.zeros(12)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    if lib().lfrv_io_mve_camera(str(meta_ini).encode(), w, h, dp(K), dp(R), dp(t), dp(P)) != 0:
        return None
    return K.reshape(3, 3), R.reshape(3, 3), t, P.reshape(3, 4)


def mve_depth(mve_name, s, scale_min, scale_max, w, h):
    """(z, dz/dx, dz/dy), each (h, w), of view s as importViews reads it, or None."""
    z, gx, gy = (np.zeros((h, w), np.float32) for _ in range(3))
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    if lib().lfrv_io_mve_depth(str(mve_name).encode(), s, scale_min, scale_max, w, h, fp(z), fp(gx), fp(gy)) != 0:
        return None
    return z, gx, gy
