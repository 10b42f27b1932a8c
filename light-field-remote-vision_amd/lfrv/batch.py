"""ctypes bindings of include/lfrv_batch.h: many novel-view solves at once."""
import ctypes

from . import LfrvError, lib as _lib

CG, PD = 0, 1


class BatchParams(ctypes.Structure):
    _fields_ = [("chunk", ctypes.c_int), ("use_graph", ctypes.c_int)]


class BatchResult(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("status", ctypes.c_int), ("iterations", ctypes.c_int),
                ("true_rel_residual", ctypes.c_double), ("done_ms", ctypes.c_double)]

    def as_dict(self):
        return dict(kind=self.kind, status=self.status, iterations=self.iterations,
                    true_rel_residual=self.true_rel_residual, done_ms=self.done_ms)


def _sig():
    L = _lib()
    vp, st = ctypes.c_void_p, ctypes.c_int
    for name, res, args in (
            ("lfrv_batch_create", st, [ctypes.POINTER(BatchParams), ctypes.POINTER(vp)]),
            ("lfrv_batch_destroy", None, [vp]),
            ("lfrv_batch_last_error", ctypes.c_char_p, [vp]),
            ("lfrv_batch_add", st, [vp, vp, st, ctypes.c_double, st, st]),
            ("lfrv_batch_run", st, [vp, ctypes.POINTER(BatchResult)])):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


class Batch:
    """An lfrv_batch over contexts the caller keeps alive."""

    def __init__(self, chunk=8, use_graph=True):
        self.L = _sig()
        self.h = ctypes.c_void_p()
        p = BatchParams(chunk, int(use_graph))
        st = self.L.lfrv_batch_create(ctypes.byref(p), ctypes.byref(self.h))
        if st != 0:
            raise LfrvError(st, "lfrv_batch_create failed")
        self.n = 0
        self.ctxs = []

    def _check(self, st):
        if st != 0:
            raise LfrvError(st, self.L.lfrv_batch_last_error(self.h).decode())

    def add_cg(self, ctx, tol=1e-5, maxit=5000):
        self._check(self.L.lfrv_batch_add(self.h, ctx._h, CG, tol, maxit, 0))
        self.ctxs.append(ctx)
        self.n += 1

    def add_pd(self, ctx, steps):
        self._check(self.L.lfrv_batch_add(self.h, ctx._h, PD, 0.0, 0, steps))
        self.ctxs.append(ctx)
        self.n += 1

    def run(self, check=True):
        res = (BatchResult * max(self.n, 1))()
        st = self.L.lfrv_batch_run(self.h, res)
        if check:
            self._check(st)
        return [res[i].as_dict() for i in range(self.n)]

    def close(self):
        if self.h:
            self.L.lfrv_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
