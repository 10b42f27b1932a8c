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
