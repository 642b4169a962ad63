"""ctypes bindings of the serving loop (sqsrv_* in include/sq_serve.h)."""
import ctypes as C

import numpy as np

from .abi import GTOBasisSet
from .engine import lib
from . import inputs


class SrvConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("slots", C.c_int), ("memBudget", C.c_double),
                ("mixedThreshold", C.c_double), ("SCFConv", C.c_double), ("SCFMax", C.c_int),
                ("threeStep", C.c_int), ("SCFDrag", C.c_double)]


class SrvStats(C.Structure):
    _fields_ = [("nstep", C.c_int), ("converged_steps", C.c_int), ("nBasis", C.c_int), ("iterations", C.c_int),
                ("arrival", C.c_double), ("start", C.c_double), ("finish", C.c_double),
                ("footprint", C.c_double), ("peak_in_flight", C.c_int), ("peak_memory", C.c_double)]


def _bind():
    L = lib()
    if getattr(L, "_srv_bound", False):
        return L
    L.sqsrv_create.argtypes = [C.POINTER(SrvConfig), C.c_int, C.POINTER(GTOBasisSet), C.POINTER(C.c_void_p)]
    L.sqsrv_submit.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_double), C.c_int,
                               C.c_int, C.c_double]
    L.sqsrv_run.argtypes = [C.c_void_p]
    L.sqsrv_result.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int),
                               C.POINTER(SrvStats)]
    L.sqsrv_destroy.argtypes = [C.c_void_p]
    L.sqsrv_errmsg.restype = C.c_char_p
    L._srv_bound = True
    return L


class BasisDB:
    """GTOBasisSet_t records of a GAMESS basis file (one per Cartesian
    component, in file order), as read_GAMESS_BasisSet produces them."""

    def __init__(self, basis):
        path = basis if "/" in str(basis) else inputs.basis_path(basis)
        db = inputs.read_gamess_basis(path)
        recs = []
        for Z in sorted(db):
            for comps, ex, co in db[Z]:
                for (l, m, n) in comps:
                    recs.append((Z, l, m, n, np.ascontiguousarray(ex, dtype=np.float64),
                                 np.ascontiguousarray(co, dtype=np.float64)))
        self.arr = (GTOBasisSet * len(recs))()
        self._keep = []
        for r, (Z, l, m, n, ex, co) in zip(self.arr, recs):
            r.name = str(basis).encode()[:255]
            r.Z, r.nContract, r.l, r.m, r.n = Z, len(ex), l, m, n
            r.exp = ex.ctypes.data_as(C.POINTER(C.c_double))
            r.coef = co.ctypes.data_as(C.POINTER(C.c_double))
            self._keep += [ex, co]
        self.size = len(recs)


class Server:
    def __init__(self, basis="631g", slots=4, mem_budget=64e9, mixed=0.0, conv=1e-6, max_iter=80,
                 three_step=True, drag=0.25, device=0):
        L = _bind()
        self.db = BasisDB(basis)
        cfg = SrvConfig(device, slots, mem_budget, mixed, conv, max_iter, int(three_step), drag)
        self._h = C.c_void_p()
        if L.sqsrv_create(C.byref(cfg), self.db.size, self.db.arr, C.byref(self._h)):
            raise RuntimeError(f"sqsrv_create failed: {L.sqsrv_errmsg().decode()}")
        self.req = []

    def submit(self, Z, xyz_bohr, charge=0, multiplicity=1, arrival=0.0):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        nstep = xyz.shape[0]
        rid = lib().sqsrv_submit(self._h, len(Z), Z.ctypes.data_as(C.POINTER(C.c_int)), nstep,
                                 xyz.ctypes.data_as(C.POINTER(C.c_double)), charge, multiplicity, arrival)
        if rid < 0:
            raise RuntimeError(f"sqsrv_submit failed: {lib().sqsrv_errmsg().decode()}")
        self.req.append(nstep)
        return rid

    def run(self):
        if lib().sqsrv_run(self._h):
            raise RuntimeError(f"sqsrv_run failed: {lib().sqsrv_errmsg().decode()}")

    def result(self, rid):
        n = self.req[rid]
        E, it = np.zeros(n), np.zeros(n, dtype=np.int32)
        st = SrvStats()
        lib().sqsrv_result(self._h, rid, E.ctypes.data_as(C.POINTER(C.c_double)),
                           it.ctypes.data_as(C.POINTER(C.c_int)), C.byref(st))
        return E, it, {k: getattr(st, k) for k, _ in SrvStats._fields_}

    def close(self):
        if self._h:
            lib().sqsrv_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve(requests, **kw):
    """Runs a list of serving_requests() dicts; returns per-request results."""
    srv = Server(**kw)
    for r in requests:
        srv.submit(r["Z"], r["xyz_bohr"], r["charge"], r["multiplicity"], r["arrival"])
    srv.run()
    out = [srv.result(i) for i in range(len(requests))]
    srv.close()
    return out
