"""ctypes mirrors of the structs that cross the C-ABI boundary.

Field order and types follow include/sq_types.h, which in turn matches
src/basis.h:21-32, src/mol.h:19-26 and src/option.h:19-180 of the
reference (non-LIBXC layout).
"""
import ctypes as C

import numpy as np


class GTOBasis(C.Structure):
    _fields_ = [
        ("nContract", C.c_int), ("l", C.c_int), ("m", C.c_int), ("n", C.c_int),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double),
        ("coef", C.POINTER(C.c_double)), ("exp", C.POINTER(C.c_double)),
        ("norm", C.POINTER(C.c_double)),
    ]


class Molecule(C.Structure):
    _fields_ = [
        ("nAtom", C.c_int), ("Q", C.c_int), ("Z", C.POINTER(C.c_int)),
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("z", C.POINTER(C.c_double)),
    ]


class GTOBasisSet(C.Structure):
    _fields_ = [
        ("name", C.c_char * 256), ("Z", C.c_int), ("nContract", C.c_int),
        ("l", C.c_int), ("m", C.c_int), ("n", C.c_int),
        ("exp", C.POINTER(C.c_double)), ("coef", C.POINTER(C.c_double)),
    ]


_S = C.c_char * 256
_I, _D = C.c_int, C.c_double


class Option(C.Structure):
    _fields_ = [
        ("molCharge", _I), ("multiplicity", _I), ("outVolumeType", _I),
        ("outDensityType", _I), ("outWhichMO", _I), ("outVolume", _I),
        ("outVolumeGrid", _I), ("outVolumeCut", _D), ("outVolumeFormat", _I),
        ("outGAUSSIAN", _I), ("outGAMESS", _I), ("outLine", _I),
        ("lineDEF", _D * 6), ("lineSPAN", _D * 2), ("lineNP", _I),
        ("SCFConv", _D), ("SCFDrag", _D), ("SCFCutoff", _D), ("SCFMax", _I),
        ("convMethod", _I), ("SCFAccuracy", _I), ("SCFGuess", _I), ("maxMem", _I),
        ("loadDMatrix", _I), ("saveDMatrix", _I), ("saveCheck", _I),
        ("saveCheckAll", _I), ("loadCheck", _I), ("opt", _I), ("optMax", _I),
        ("restricted", _I), ("method", _I), ("whichHF2xDFT", _I),
        ("whichGrid", _I), ("nradial", _I), ("nlabedev", _I), ("nradialInit", _I),
        ("nlabedevInit", _I), ("force", _I), ("hessian", _I),
        ("DMatrixFile", _S), ("CheckFile", _S), ("nCPU", _I), ("prefixStr", _S),
        ("MECP", _I), ("mecpMax", _I), ("mecpMA", _I), ("mecpMB", _I),
        ("DMatrixFileA", _S), ("DMatrixFileB", _S), ("CheckFileA", _S),
        ("CheckFileB", _S), ("gaussEXE", _S), ("gaussINA", _S), ("gaussINB", _S),
        ("excite", _I), ("Ex", _D), ("Ey", _D), ("Ez", _D), ("qmd", _I),
        ("initTemp", _D), ("keepTemp", _D), ("initVel", _S), ("traj", _S),
        ("qmdFreq", _D), ("tStep", _D), ("maxQMDStep", _I),
    ]


CONVMETHOD_DAMPING, CONVMETHOD_DIIS2, CONVMETHOD_DIIS3, CONVMETHOD_DIIS4 = 0, 1, 2, 3
SCFACCURACY_3STEP, SCFACCURACY_1STEP = 0, 1
SCFGUESS_DIAG, SCFGUESS_CORE, SCFGUESS_CHECK, SCFGUESS_CACB = 0, 1, 2, 3


def default_option(**kw):
    """Option struct with the reference defaults (src/option.c:495-583)."""
    o = Option()
    o.restricted = 1
    o.SCFGuess = SCFGUESS_DIAG
    o.SCFConv = 1.0e-6
    o.SCFCutoff = 1.0e-15
    o.SCFDrag = 0.25
    o.SCFMax = 80
    o.convMethod = CONVMETHOD_DIIS4
    o.SCFAccuracy = SCFACCURACY_3STEP
    o.maxMem = 250
    o.nCPU = 1
    o.optMax = 30
    o.tStep = 0.001
    o.maxQMDStep = 25
    o.outVolumeGrid = 10
    o.outVolumeCut = 1.0e-4
    o.lineNP = 13
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Basis:
    """Owns a GTOBasis_t array built in Python (keeps the numpy buffers alive)."""

    def __init__(self, funcs):
        # funcs: list of (l, m, n, (x, y, z), exps, coefs, norms)
        self.n = len(funcs)
        self.arr = (GTOBasis * max(self.n, 1))()
        self._keep = []
        for i, (l, m, n, c, ex, co, no) in enumerate(funcs):
            ex = np.ascontiguousarray(ex, dtype=np.float64)
            co = np.ascontiguousarray(co, dtype=np.float64)
            no = np.ascontiguousarray(no, dtype=np.float64)
            self._keep += [ex, co, no]
            g = self.arr[i]
            g.nContract, g.l, g.m, g.n = len(ex), l, m, n
            g.x0, g.y0, g.z0 = c
            g.exp = ex.ctypes.data_as(C.POINTER(C.c_double))
            g.coef = co.ctypes.data_as(C.POINTER(C.c_double))
            g.norm = no.ctypes.data_as(C.POINTER(C.c_double))

    @property
    def ptr(self):
        return C.cast(self.arr, C.POINTER(GTOBasis))


def basis_from_ptr(ptr, n):
    """Python list-of-tuples view of a C GTOBasis_t array (copies the data)."""
    out = []
    for i in range(n):
        g = ptr[i]
        k = g.nContract
        out.append((g.l, g.m, g.n, (g.x0, g.y0, g.z0),
                    np.ctypeslib.as_array(g.exp, (k,)).copy(),
                    np.ctypeslib.as_array(g.coef, (k,)).copy(),
                    np.ctypeslib.as_array(g.norm, (k,)).copy()))
    return out


class Mol:
    """Owns a Molecule_t built in Python; coordinates in bohr."""

    def __init__(self, Z, xyz_bohr, charge=0):
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self.x = np.ascontiguousarray(xyz[:, 0])
        self.y = np.ascontiguousarray(xyz[:, 1])
        self.z = np.ascontiguousarray(xyz[:, 2])
        self.s = Molecule()
        self.s.nAtom = len(self.Z)
        self.s.Q = charge
        self.s.Z = self.Z.ctypes.data_as(C.POINTER(C.c_int))
        self.s.x = self.x.ctypes.data_as(C.POINTER(C.c_double))
        self.s.y = self.y.ctypes.data_as(C.POINTER(C.c_double))
        self.s.z = self.z.ctypes.data_as(C.POINTER(C.c_double))

    @property
    def ptr(self):
        return C.pointer(self.s)

    @property
    def n_electrons(self):
        return int(self.Z.sum()) - self.s.Q


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))
