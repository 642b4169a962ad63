"""ctypes bindings of include/sq_jk.h (context API and drop-in entry points)."""
import ctypes as C
import os

import numpy as np

from .abi import GTOBasis, Option, dptr

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PD = C.POINTER(C.c_double)
PG = C.POINTER(GTOBasis)
_lib = None


def library_path():
    return os.path.join(PKG, "lib", "libsqamd.so")


class SqjkStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("quartets", C.c_longlong),
                ("quartets_fp32", C.c_longlong), ("launches", C.c_int)]


class Mp2Info(C.Structure):
    """sqjk_mp2_info (include/sq_mp2.h)."""
    _fields_ = [("e2", C.c_double), ("store_ms", C.c_double), ("half1_ms", C.c_double), ("half2_ms", C.c_double),
                ("half1_flop", C.c_double), ("half2_flop", C.c_double), ("quartets", C.c_longlong),
                ("device_bytes", C.c_longlong)]


def lib():
    """Load lib/libsqamd.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(there is no CPU fallback)")
    L = C.CDLL(path)
    L.sqjk_create.argtypes = [C.c_int, PG, PD, C.c_int, C.POINTER(C.c_void_p)]
    L.sqjk_build.argtypes = [C.c_void_p, PD, PD, C.c_double, C.c_double, C.c_int, PD, PD, PD]
    L.sqjk_schwarz.argtypes = [C.c_void_p, PD]
    L.sqjk_set_mixed.argtypes = [C.c_void_p, C.c_double]
    L.sqjk_last_stats.argtypes = [C.c_void_p, C.POINTER(SqjkStats)]
    L.sqjk_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.sqjk_destroy.argtypes = [C.c_void_p]
    L.sqjk_set_profile.argtypes = [C.c_void_p, C.c_int]
    L.sqjk_class_times.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.sqjk_class_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong),
                                   C.POINTER(C.c_longlong)]
    PLL = C.POINTER(C.c_longlong)
    L.sqjk_class_work.argtypes = [C.c_void_p, PLL, PLL, PLL, PLL]
    L.sqjk_errmsg.restype = C.c_char_p
    L.GTO_JK_Matrix_Quartet_Parallel.argtypes = [C.c_int, C.c_int, PD, PD, PG, PD, C.c_double,
                                                 PD, PD, PD, C.POINTER(Option)]
    L.uhf_getGMatrix.argtypes = [C.c_int, PG, PD, C.c_double, PD, PD, PD, PD, PD, C.POINTER(Option)]
    L.create_Schwarz.argtypes = [C.c_int, PG]
    L.create_Schwarz.restype = PD
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {lib().sqjk_errmsg().decode()}")


class JKEngine:
    """One molecule's J/K context on the device (sqjk_create/sqjk_build)."""

    def __init__(self, basis_ptr, n, schwarz=None, device=0):
        L = lib()
        self.n = n
        self._h = C.c_void_p()
        sch = None if schwarz is None else np.ascontiguousarray(schwarz, dtype=np.float64)
        _check(L.sqjk_create(n, basis_ptr, dptr(sch) if sch is not None else None, device,
                             C.byref(self._h)), "sqjk_create")

    def build(self, PA, PB=None, fixed_cutoff=0.0, scf_cutoff=0.0, restricted=True, out=None):
        """Returns (JT, KA, KB) accumulated onto `out` (zeros by default)."""
        n = self.n
        PA = np.ascontiguousarray(PA, dtype=np.float64)
        PB = PA if PB is None else np.ascontiguousarray(PB, dtype=np.float64)
        if out is None:
            out = (np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n)))
        J, KA, KB = out
        _check(lib().sqjk_build(self._h, dptr(PA), dptr(PB), fixed_cutoff, scf_cutoff,
                                int(bool(restricted)), dptr(J), dptr(KA), dptr(KB)), "sqjk_build")
        return J, KA, KB

    def set_tuning(self, rows=None, split=None, grid=None):
        """Work-plan tuning tables (6 x 6 ints each, see sqjk_set_tuning);
        all None restores the defaults."""
        L = lib()
        P32 = C.POINTER(C.c_int)
        L.sqjk_set_tuning.argtypes = [C.c_void_p, P32, P32, P32]
        arr = [None if t is None else np.ascontiguousarray(np.asarray(t).reshape(36), dtype=np.int32)
               for t in (rows, split, grid)]
        _check(L.sqjk_set_tuning(self._h, *[None if a is None else a.ctypes.data_as(P32) for a in arr]),
               "sqjk_set_tuning")

    def set_mixed(self, threshold):
        """Quartets with Schwarz product below `threshold` run in FP32 (0: all FP64)."""
        _check(lib().sqjk_set_mixed(self._h, float(threshold)), "sqjk_set_mixed")

    def schwarz(self):
        s = np.zeros((self.n, self.n))
        _check(lib().sqjk_schwarz(self._h, dptr(s)), "sqjk_schwarz")
        return s

    def stats(self):
        st = SqjkStats()
        _check(lib().sqjk_last_stats(self._h, C.byref(st)), "sqjk_last_stats")
        return {k: getattr(st, k) for k, _ in SqjkStats._fields_}

    def class_stats(self):
        """[(la,lb,lc,ld), quartets, primitive quartets] of the last build."""
        lq = (C.c_int * 84)()
        q = (C.c_longlong * 21)()
        p = (C.c_longlong * 21)()
        k = lib().sqjk_class_stats(self._h, lq, q, p)
        return [((lq[4 * i], lq[4 * i + 1], lq[4 * i + 2], lq[4 * i + 3]), q[i], p[i]) for i in range(k)]

    def class_work(self):
        """Surviving work of the last build per class: [(la,lb,lc,ld), quartets,
        primitive quartets, quartets in FP32, primitive quartets in FP32],
        after the density screen and the primitive cut-off."""
        q, p, q32, p32 = [(C.c_longlong * 21)() for _ in range(4)]
        k = lib().sqjk_class_work(self._h, q, p, q32, p32)
        if k < 0:
            _check(k, "sqjk_class_work")
        lq = [c[0] for c in self.class_stats()]
        return [(lq[i], q[i], p[i], q32[i], p32[i]) for i in range(k)]

    def set_profile(self, on=True):
        _check(lib().sqjk_set_profile(self._h, int(on)), "sqjk_set_profile")

    def grad_eri(self, PA, PB, basis2atom, natom, kfactor=1.0, cutoff=1e-10):
        """Two-electron energy gradient per atom, [natom, 3] (sqjk_grad_eri)."""
        L = lib()
        P32 = C.POINTER(C.c_int)
        L.sqjk_grad_eri.argtypes = [C.c_void_p, PD, PD, C.c_double, C.c_double, P32, C.c_int, PD]
        PA = np.ascontiguousarray(PA, dtype=np.float64)
        PB = np.ascontiguousarray(PB, dtype=np.float64)
        b2a = np.ascontiguousarray(basis2atom, dtype=np.int32)
        G = np.zeros((natom, 3))
        _check(L.sqjk_grad_eri(self._h, dptr(PA), dptr(PB), float(kfactor), float(cutoff),
                               b2a.ctypes.data_as(P32), int(natom), dptr(G)), "sqjk_grad_eri")
        return G

    def eri_pairs(self, cutoff=0.0):
        """Dense pair matrix V[pair(p,q), pair(r,s)] = (pq|rs) of every shell
        quartet above the Schwarz cut-off (sqjk_eri_pairs)."""
        L = lib()
        L.sqjk_eri_pairs.argtypes = [C.c_void_p, C.c_double, PD]
        m = self.n * (self.n + 1) // 2
        V = np.zeros((m, m))
        _check(L.sqjk_eri_pairs(self._h, float(cutoff), dptr(V)), "sqjk_eri_pairs")
        return V

    def mp2_rhf(self, nocc, ncore, e, Cm, cutoff=0.0):
        """Closed-shell MP2 correlation energy and timings (sqjk_mp2_rhf);
        Cm: orbital-major coefficients C[k*n + i]."""
        L = lib()
        L.sqjk_mp2_rhf.argtypes = [C.c_void_p, C.c_int, C.c_int, PD, PD, C.c_double, C.POINTER(Mp2Info)]
        e = np.ascontiguousarray(e, dtype=np.float64)
        Cm = np.ascontiguousarray(Cm, dtype=np.float64)
        info = Mp2Info()
        _check(L.sqjk_mp2_rhf(self._h, int(nocc), int(ncore), dptr(e), dptr(Cm), float(cutoff), C.byref(info)),
               "sqjk_mp2_rhf")
        return {f: getattr(info, f) for f, _ in Mp2Info._fields_}

    def mp2_uhf(self, nea, neb, ncore, eA, eB, CA, CB, cutoff=0.0):
        """Unrestricted MP2 correlation energy and timings (sqjk_mp2_uhf)."""
        L = lib()
        L.sqjk_mp2_uhf.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, PD, PD, PD, PD, C.c_double,
                                   C.POINTER(Mp2Info)]
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (eA, eB, CA, CB)]
        info = Mp2Info()
        _check(L.sqjk_mp2_uhf(self._h, int(nea), int(neb), int(ncore), *[dptr(x) for x in a], float(cutoff),
                              C.byref(info)), "sqjk_mp2_uhf")
        return {f: getattr(info, f) for f, _ in Mp2Info._fields_}

    def class_times(self):
        """Kernel ms per class of the last build (profile mode), class_stats order."""
        t = (C.c_double * 21)()
        k = lib().sqjk_class_times(self._h, t)
        return [t[i] for i in range(k)]

    def dims(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().sqjk_dims(self._h, C.byref(a), C.byref(b), C.byref(c)), "sqjk_dims")
        return a.value, b.value, c.value

    def close(self):
        if self._h:
            lib().sqjk_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grad_one_electron(bas, mol, PT, QT, basis2atom, device=0):
    """One-electron energy gradient per atom, [natom, 3] (sqjk_grad_one_electron)."""
    from .abi import Molecule
    L = lib()
    P32 = C.POINTER(C.c_int)
    L.sqjk_grad_one_electron.argtypes = [C.c_int, PG, C.POINTER(Molecule), PD, PD, P32, C.c_int, PD]
    PT = np.ascontiguousarray(PT, dtype=np.float64)
    QT = np.ascontiguousarray(QT, dtype=np.float64)
    b2a = np.ascontiguousarray(basis2atom, dtype=np.int32)
    na = mol.ptr.contents.nAtom
    G = np.zeros((na, 3))
    _check(L.sqjk_grad_one_electron(bas.n, bas.ptr, mol.ptr, dptr(PT), dptr(QT), b2a.ctypes.data_as(P32),
                                    int(device), dptr(G)), "sqjk_grad_one_electron")
    return G


def basis_to_atom(mol, bas):
    """Atom of each basis function (the first atom at its centre)."""
    m, g = mol.ptr.contents, bas.ptr
    out = np.empty(bas.n, dtype=np.int32)
    for p in range(bas.n):
        out[p] = next(i for i in range(m.nAtom) if (g[p].x0, g[p].y0, g[p].z0) == (m.x[i], m.y[i], m.z[i]))
    return out


def uhf_force(bas, mol, nEA, nEB, CA, CB, eA, eB, opt):
    """The uhf_force drop-in: forces -dE/dR per atom, [natom, 3]."""
    from .abi import Molecule
    L = lib()
    L.uhf_force.argtypes = [C.c_int, PG, C.POINTER(Molecule), C.c_int, C.c_int, PD, PD, PD, PD,
                            C.POINTER(Option), PD, PD, PD]
    CA, CB, eA, eB = (np.ascontiguousarray(a, dtype=np.float64) for a in (CA, CB, eA, eB))
    na = mol.ptr.contents.nAtom
    F = np.zeros((3, na))
    L.uhf_force(bas.n, bas.ptr, mol.ptr, nEA, nEB, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt),
                dptr(F[0]), dptr(F[1]), dptr(F[2]))
    return F.T.copy()


def mp2_rhf_aqij_parallel(bas, mol, nocc, e, Cm, opt):
    """The mp2_rhf_aqij_Parallel drop-in: frozen-core RHF MP2 energy
    correction (prints the reference's report)."""
    from .abi import Molecule
    L = lib()
    L.mp2_rhf_aqij_Parallel.argtypes = [C.c_int, C.c_int, PD, PD, PG, C.POINTER(Molecule), C.POINTER(Option)]
    L.mp2_rhf_aqij_Parallel.restype = C.c_double
    e = np.ascontiguousarray(e, dtype=np.float64)
    Cm = np.ascontiguousarray(Cm, dtype=np.float64)
    return L.mp2_rhf_aqij_Parallel(bas.n, int(nocc), dptr(e), dptr(Cm), bas.ptr, mol.ptr, C.byref(opt))


def frozen_core(mol):
    """getFrozenCore_RHF of the drop-in library."""
    from .abi import Molecule
    L = lib()
    L.getFrozenCore_RHF.argtypes = [C.POINTER(Molecule)]
    return L.getFrozenCore_RHF(mol.ptr)


def mp2_uhf_aqij(bas, mol, nea, neb, eA, eB, CA, CB, opt):
    """The mp2_uhf_aqij drop-in: frozen-core unrestricted MP2 energy
    correction (prints the reference's report)."""
    from .abi import Molecule
    L = lib()
    L.mp2_uhf_aqij.argtypes = [C.c_int, C.c_int, C.c_int, PD, PD, PD, PD, PG, C.POINTER(Molecule),
                               C.POINTER(Option)]
    L.mp2_uhf_aqij.restype = C.c_double
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (eA, eB, CA, CB)]
    return L.mp2_uhf_aqij(bas.n, int(nea), int(neb), *[dptr(x) for x in a], bas.ptr, mol.ptr, C.byref(opt))
