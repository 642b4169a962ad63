"""ctypes bindings of the batched J/K builds (sqjk_batch_* in include/sq_jk.h)."""
import ctypes as C

import numpy as np

from .abi import dptr
from .engine import lib

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


def _bind():
    L = lib()
    if getattr(L, "_batch_bound", False):
        return L
    L.sqjk_batch_create.argtypes = [C.c_int, C.POINTER(C.c_void_p), PI, C.POINTER(C.c_void_p)]
    L.sqjk_batch_set_density.argtypes = [C.c_void_p, C.c_int, PD, PD, C.c_int]
    L.sqjk_batch_set_cutoffs.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
    L.sqjk_batch_set_mixed.argtypes = [C.c_void_p, C.c_double]
    L.sqjk_batch_set_tuning.argtypes = [C.c_void_p, PI, PI, PI]
    L.sqjk_batch_set_profile.argtypes = [C.c_void_p, C.c_int]
    L.sqjk_batch_build.argtypes = [C.c_void_p, C.c_int]
    L.sqjk_batch_result.argtypes = [C.c_void_p, C.c_int, PD, PD, PD, C.c_int]
    L.sqjk_batch_stats.argtypes = [C.c_void_p, PD, C.POINTER(C.c_longlong), PI, PD]
    L.sqjk_batch_destroy.argtypes = [C.c_void_p]
    L.sqjk_set_tuning.argtypes = [C.c_void_p, PI, PI, PI]
    L.sqjk_get_tuning.argtypes = [C.c_void_p, PI, PI, PI]
    L._batch_bound = True
    return L


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {lib().sqjk_errmsg().decode()}")


def _table(t):
    if t is None:
        return None
    a = np.ascontiguousarray(np.asarray(t, dtype=np.int32).reshape(36))
    return a


class JKBatch:
    """J/K builds of many molecules (JKEngine contexts, kept alive by the
    caller) as one set of class launches, optionally replayed from a HIP
    graph."""

    def __init__(self, engines, restricted):
        L = _bind()
        self.engines = list(engines)
        self.n = [e.n for e in self.engines]
        hs = (C.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        r = np.ascontiguousarray(np.asarray(restricted, dtype=np.int32))
        self._r = r
        self._h = C.c_void_p()
        _check(L.sqjk_batch_create(len(self.engines), hs, r.ctypes.data_as(PI), C.byref(self._h)),
               "sqjk_batch_create")

    def set_density(self, i, PA, PB=None):
        PA = np.ascontiguousarray(PA, dtype=np.float64)
        PB = PA if PB is None else np.ascontiguousarray(PB, dtype=np.float64)
        _check(lib().sqjk_batch_set_density(self._h, i, dptr(PA), dptr(PB), 0), "sqjk_batch_set_density")

    def set_cutoffs(self, fixed, scf, i=-1):
        _check(lib().sqjk_batch_set_cutoffs(self._h, i, fixed, scf), "sqjk_batch_set_cutoffs")

    def set_mixed(self, thr):
        _check(lib().sqjk_batch_set_mixed(self._h, float(thr)), "sqjk_batch_set_mixed")

    def set_tuning(self, rows=None, split=None, grid=None):
        a, b, c = _table(rows), _table(split), _table(grid)
        p = [x.ctypes.data_as(PI) if x is not None else None for x in (a, b, c)]
        _check(lib().sqjk_batch_set_tuning(self._h, *p), "sqjk_batch_set_tuning")

    def set_profile(self, on=True):
        _check(lib().sqjk_batch_set_profile(self._h, int(on)), "sqjk_batch_set_profile")

    def build(self, graph=True):
        _check(lib().sqjk_batch_build(self._h, int(graph)), "sqjk_batch_build")

    def result(self, i):
        n = self.n[i]
        J, KA, KB = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        _check(lib().sqjk_batch_result(self._h, i, dptr(J), dptr(KA), dptr(KB), 0), "sqjk_batch_result")
        return J, KA, KB

    def stats(self):
        ms, q, nl = C.c_double(), C.c_longlong(), C.c_int()
        cls = (C.c_double * 21)()
        _check(lib().sqjk_batch_stats(self._h, C.byref(ms), C.byref(q), C.byref(nl), cls), "sqjk_batch_stats")
        return {"kernel_ms": ms.value, "quartets": q.value, "launches": nl.value, "class_ms": list(cls)}

    def close(self):
        if self._h:
            lib().sqjk_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
