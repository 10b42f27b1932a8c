"""ctypes bindings of include/lfrv_cnn.h (the warm-start U-Net) and of the
device scene renderer, with bf16 bit conversions and raw device buffers
through the library's allocator.  Every numeric call goes to liblfrv.so."""
import ctypes

import numpy as np

from . import LfrvError, SynthDesc, lib as _lfrv_lib


class _Lfrv:
    """The parts of the parent module used here (avoids a circular import)."""

    @staticmethod
    def lib():
        return _lfrv_lib()

    LfrvError = LfrvError


lfrv = _Lfrv

_vp = ctypes.c_void_p
_st = ctypes.c_int


def to_bf16(a):
    """float32 -> bf16 bits, round to nearest even (the library's f2bf)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


class Raw:
    """A device buffer of raw bytes (library allocator)."""

    def __init__(self, arr_or_nbytes):
        L = lfrv.lib()
        self.ptr = ctypes.c_void_p()
        if isinstance(arr_or_nbytes, np.ndarray):
            a = np.ascontiguousarray(arr_or_nbytes)
            self.nbytes = a.nbytes
            assert L.lfrv_dev_alloc(self.nbytes, ctypes.byref(self.ptr)) == 0
            assert L.lfrv_memcpy(self.ptr, a.ctypes.data, self.nbytes) == 0
        else:
            self.nbytes = int(arr_or_nbytes)
            assert L.lfrv_dev_alloc(self.nbytes, ctypes.byref(self.ptr)) == 0
            z = np.zeros(self.nbytes, np.uint8)
            assert L.lfrv_memcpy(self.ptr, z.ctypes.data, self.nbytes) == 0

    def get(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        assert lfrv.lib().lfrv_memcpy(out.ctypes.data, self.ptr, self.nbytes) == 0
        return out

    def free(self):
        if self.ptr:
            lfrv.lib().lfrv_dev_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CnnParams(ctypes.Structure):
    _fields_ = [("in_channels", ctypes.c_int), ("levels", ctypes.c_int), ("base_channels", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("max_h", ctypes.c_int), ("max_w", ctypes.c_int),
                ("seed", ctypes.c_uint64)]


class WarmStats(ctypes.Structure):
    _fields_ = [("energy_plain", ctypes.c_double), ("energy_net", ctypes.c_double), ("used_net", ctypes.c_int),
                ("ms", ctypes.c_double)]


def cnn_lib():
    L = lfrv.lib()
    sig = {
        "lfrv_cnn_create": (_st, [ctypes.POINTER(CnnParams), ctypes.POINTER(_vp)]),
        "lfrv_cnn_destroy": (None, [_vp]),
        "lfrv_cnn_last_error": (ctypes.c_char_p, [_vp]),
        "lfrv_cnn_num_params": (_st, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
        "lfrv_cnn_get_params": (_st, [_vp, _vp]),
        "lfrv_cnn_set_params": (_st, [_vp, _vp]),
        "lfrv_cnn_get_grads": (_st, [_vp, _vp]),
        "lfrv_cnn_set_batch": (_st, [_vp, _st, _st, _st, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                     ctypes.POINTER(_vp)]),
        "lfrv_cnn_set_batch_host": (_st, [_vp, _st, _st, _st, _vp, _vp, _vp]),
        "lfrv_cnn_forward": (_st, [_vp, _vp]),
        "lfrv_cnn_loss_grad": (_st, [_vp, ctypes.POINTER(ctypes.c_double)]),
        "lfrv_cnn_train_step": (_st, [_vp, ctypes.c_float, ctypes.POINTER(ctypes.c_double)]),
        "lfrv_cnn_forward_flops": (_st, [_vp, ctypes.POINTER(ctypes.c_double)]),
        "lfrv_cnn_set_timing": (_st, [_vp, _st]),
        "lfrv_cnn_get_timing": (_st, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_double)]),
        "lfrv_op_conv3x3": (_st, [_st, _st, _st, _st, _st, _vp, _vp, _vp, _vp, _vp]),
        "lfrv_op_conv3x3_wgrad": (_st, [_st, _st, _st, _st, _st, _vp, _vp, _vp, _vp]),
        "lfrv_warped_stack": (_st, [_vp, _vp, _st]),
        "lfrv_warm_start": (_st, [_vp, _vp, ctypes.POINTER(WarmStats)]),
        "lfrv_synth_generate_device": (_st, [ctypes.POINTER(SynthDesc), _vp, _vp, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double), _vp, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


class Cnn:
    """An lfrv_cnn."""

    def __init__(self, in_channels, levels=4, base=32, max_batch=1, max_h=64, max_w=64, seed=1):
        self.L = cnn_lib()
        self.p = CnnParams(in_channels, levels, base, max_batch, max_h, max_w, seed)
        self.h = ctypes.c_void_p()
        st = self.L.lfrv_cnn_create(ctypes.byref(self.p), ctypes.byref(self.h))
        if st != 0:
            raise lfrv.LfrvError(st, "lfrv_cnn_create failed")
        self.cpad = (in_channels + 31) // 32 * 32

    def check(self, st):
        if st != 0:
            raise lfrv.LfrvError(st, self.L.lfrv_cnn_last_error(self.h).decode())

    def nparams(self):
        n = ctypes.c_size_t()
        self.check(self.L.lfrv_cnn_num_params(self.h, ctypes.byref(n)))
        return n.value

    def get_params(self):
        out = np.empty(self.nparams(), np.float32)
        self.check(self.L.lfrv_cnn_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self.check(self.L.lfrv_cnn_set_params(self.h, w.ctypes.data))

    def get_grads(self):
        out = np.empty(self.nparams(), np.float32)
        self.check(self.L.lfrv_cnn_get_grads(self.h, out.ctypes.data))
        return out

    def set_batch_host(self, x, base, target=None):
        x = np.ascontiguousarray(x, np.float32)
        base = np.ascontiguousarray(base, np.float32)
        t = None if target is None else np.ascontiguousarray(target, np.float32)
        B, H, W = x.shape[0], x.shape[1], x.shape[2]
        self.check(self.L.lfrv_cnn_set_batch_host(self.h, B, H, W, x.ctypes.data, base.ctypes.data,
                                                  None if t is None else t.ctypes.data))
        self.shape = (B, H, W)

    def set_batch(self, B, H, W, xs, bases, targets=None):
        """Device pointers (ints or c_void_p) of each image's input, base and target."""
        arr = lambda ps: (ctypes.c_void_p * B)(*[p.value if isinstance(p, ctypes.c_void_p) else p for p in ps])  # noqa: E731
        self.check(self.L.lfrv_cnn_set_batch(self.h, B, H, W, arr(xs), arr(bases),
                                             None if targets is None else arr(targets)))
        self.shape = (B, H, W)

    def forward(self):
        B, H, W = self.shape
        out = np.empty((B, 3, H, W), np.float32)
        self.check(self.L.lfrv_cnn_forward(self.h, out.ctypes.data))
        return out

    def loss_grad(self):
        v = ctypes.c_double()
        self.check(self.L.lfrv_cnn_loss_grad(self.h, ctypes.byref(v)))
        return v.value

    def train_step(self, lr):
        v = ctypes.c_double()
        self.check(self.L.lfrv_cnn_train_step(self.h, lr, ctypes.byref(v)))
        return v.value

    def forward_flops(self):
        v = ctypes.c_double()
        self.check(self.L.lfrv_cnn_forward_flops(self.h, ctypes.byref(v)))
        return v.value

    def set_timing(self, on):
        self.check(self.L.lfrv_cnn_set_timing(self.h, int(on)))

    def get_timing(self):
        ms, n, fl = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        self.check(self.L.lfrv_cnn_get_timing(self.h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(fl)))
        return ms.value, n.value, fl.value

    def close(self):
        if self.h:
            self.L.lfrv_cnn_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def layer_slices(levels, base, cpad):
    """(name, W slice, b slice) of each convolution in the parameter vector."""
    ch = [min(base << l, 256) for l in range(levels)]
    convs = []
    cin = cpad
    for l in range(levels):
        convs += [(f"enc{l}a", cin, ch[l]), (f"enc{l}b", ch[l], ch[l])]
        cin = ch[l]
    for l in range(levels - 2, -1, -1):
        convs += [(f"dec{l}a", ch[l + 1], ch[l]), (f"dec{l}b", ch[l], ch[l])]
    convs.append(("head", ch[0], 32))
    out, off = [], 0
    for name, ci, co in convs:
        out.append((name, slice(off, off + co * 9 * ci), slice(off + co * 9 * ci, off + co * 9 * ci + co), ci, co))
        off += co * 9 * ci + co
    return out
