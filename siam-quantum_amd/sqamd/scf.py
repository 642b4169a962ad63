"""ctypes bindings of the device-resident SCF (sqscf_* in include/sq_scf.h).

DeviceSCF keeps one molecule's SCF state on the device across solves: a
later geometry of the same basis keeps the J/K context and can start from
the previous orbitals (the reference's geometry-step guess,
src/qmd.c:585-589).
"""
import ctypes as C

import numpy as np

from .abi import GTOBasis, CONVMETHOD_DIIS4, dptr
from .engine import lib

PD = C.POINTER(C.c_double)
ONITER = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double)


class SCFParams(C.Structure):
    _fields_ = [("restricted", C.c_int), ("nEA", C.c_int), ("nEB", C.c_int), ("SCFConv", C.c_double),
                ("SCFMax", C.c_int), ("SCFDrag", C.c_double), ("threeStep", C.c_int), ("convMethod", C.c_int),
                ("Econst", C.c_double), ("mixedThreshold", C.c_double), ("eigenEveryIter", C.c_int),
                ("onIter", ONITER), ("user", C.c_void_p)]


class SCFResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("dE", C.c_double), ("rmsDP", C.c_double), ("realCutoff", C.c_double),
                ("seconds", C.c_double), ("iterations", C.c_int), ("converged", C.c_int)]


def _bind():
    L = lib()
    if getattr(L, "_scf_bound", False):
        return L
    L.sqscf_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    L.sqscf_set_system.argtypes = [C.c_void_p, C.c_int, C.POINTER(GTOBasis), PD, PD]
    L.sqscf_set_density.argtypes = [C.c_void_p, PD, PD]
    L.sqscf_guess_from_orbitals.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.sqscf_run.argtypes = [C.c_void_p, C.POINTER(SCFParams), C.POINTER(SCFResult)]
    L.sqscf_orbitals.argtypes = [C.c_void_p, PD, PD, PD, PD]
    L.sqscf_density.argtypes = [C.c_void_p, PD, PD]
    L.sqscf_destroy.argtypes = [C.c_void_p]
    L.sqscf_errmsg.restype = C.c_char_p
    L.sqjk_one_electron.argtypes = [C.c_int, C.POINTER(GTOBasis), C.c_void_p, C.c_int, PD, PD, PD]
    L.nuclei_coulomb.argtypes = [C.c_void_p]
    L.nuclei_coulomb.restype = C.c_double
    L._scf_bound = True
    return L


def one_electron(bas, mol):
    """(H = T + V, S) on the device (sqjk_one_electron)."""
    L = _bind()
    n = bas.n
    T, V, S = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    rc = L.sqjk_one_electron(n, bas.ptr, C.cast(mol.ptr, C.c_void_p), 0, dptr(T), dptr(V), dptr(S))
    if rc:
        raise RuntimeError(f"sqjk_one_electron failed ({rc}): {L.sqjk_errmsg().decode()}")
    return T + V, S


class DeviceSCF:
    def __init__(self, device=0):
        L = _bind()
        self._h = C.c_void_p()
        self._check(L.sqscf_create(device, C.byref(self._h)), "sqscf_create")
        self.n = 0

    @staticmethod
    def _check(rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc}): {lib().sqscf_errmsg().decode()}")

    def set_system(self, bas, mol):
        """Basis and geometry; H and S from the device one-electron kernel."""
        H, S = one_electron(bas, mol)
        self.n = bas.n
        self.mol = mol
        self._check(lib().sqscf_set_system(self._h, bas.n, bas.ptr, dptr(H), dptr(S)), "sqscf_set_system")
        return H, S

    def set_density(self, PA, PB):
        PA = np.ascontiguousarray(PA, dtype=np.float64)
        PB = np.ascontiguousarray(PB, dtype=np.float64)
        self._check(lib().sqscf_set_density(self._h, dptr(PA), dptr(PB)), "sqscf_set_density")

    def diag_guess(self, S, nEA, nEB):
        """The reference's DIAG guess, P_ii = nE / tr S (src/uhf.c:1384-1399)."""
        tr = np.trace(S)
        self.set_density(np.eye(self.n) * nEA / tr, np.eye(self.n) * nEB / tr)

    def guess_from_orbitals(self, nEA, nEB):
        self._check(lib().sqscf_guess_from_orbitals(self._h, nEA, nEB), "sqscf_guess_from_orbitals")

    def run(self, nEA, nEB, restricted=True, conv=1e-8, max_iter=80, drag=0.25, three_step=True,
            method=CONVMETHOD_DIIS4, mixed=0.0):
        p = SCFParams()
        p.restricted, p.nEA, p.nEB = int(restricted), nEA, nEB
        p.SCFConv, p.SCFMax, p.SCFDrag = conv, max_iter, drag
        p.threeStep, p.convMethod = int(three_step), method
        p.Econst = lib().nuclei_coulomb(C.cast(self.mol.ptr, C.c_void_p))
        p.mixedThreshold = mixed
        r = SCFResult()
        self._check(lib().sqscf_run(self._h, C.byref(p), C.byref(r)), "sqscf_run")
        return {k: getattr(r, k) for k, _ in SCFResult._fields_}

    def orbitals(self):
        n = self.n
        CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
        self._check(lib().sqscf_orbitals(self._h, dptr(CA), dptr(CB), dptr(eA), dptr(eB)), "sqscf_orbitals")
        return CA, CB, eA, eB

    def density(self):
        n = self.n
        PA, PB = np.zeros((n, n)), np.zeros((n, n))
        self._check(lib().sqscf_density(self._h, dptr(PA), dptr(PB)), "sqscf_density")
        return PA, PB

    def close(self):
        if self._h:
            lib().sqscf_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
