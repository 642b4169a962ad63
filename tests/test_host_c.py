"""Host-side C functions of the product (no GPU needed): Boys, one-electron
integrals, host ERI, basis/molecule readers — checked against mpmath and the
oracle restatement of src/int.c, src/basis.c, src/mol.c."""
import ctypes as C
import os

import numpy as np
import pytest

from sqamd import GTOBasis, GTOBasisSet, Molecule, basis_from_ptr, dptr, lib
from sqamd.inputs import basis_path

HERE = os.path.dirname(os.path.abspath(__file__))
D, I = C.c_double, C.c_int


@pytest.fixture(scope="module")
def L():
    L = lib()
    L.fgamma_set.argtypes = [I, D, C.POINTER(D)]
    for f in ("overlap", "kinetic"):
        getattr(L, f).argtypes = [D, I, I, I, D, D, D] * 2
        getattr(L, f).restype = D
    L.moment.argtypes = [D, I, I, I, D, D, D] * 2 + [I, I, I, D, D, D]
    L.moment.restype = D
    L.nai.argtypes = [D, D, D, D, I, I, I, D] * 2 + [D, D, D]
    L.nai.restype = D
    L.eri.argtypes = [D, D, D, D, I, I, I, D] * 4
    L.eri.restype = D
    return L


def test_boys_vs_mpmath(L):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    for T in [0.0, 1e-6, 3e-4, 0.2, 2.5, 9.0, 29.9, 30.1, 44.0, 150.0]:
        F = np.zeros(17)
        L.fgamma_set(16, T, dptr(F))
        for m in range(17):
            ex = (mpmath.mpf(1) / (2 * m + 1) if T == 0 else
                  mpmath.quad(lambda t: t ** (2 * m) * mpmath.e ** (-T * t * t), [0, 1]))
            assert abs((F[m] - ex) / ex) < 2e-15, (T, m)


def _rand_prims(rng, k, lmax):
    out = []
    for _ in range(k):
        l = [int(v) for v in rng.integers(0, lmax + 1, 3)]
        while sum(l) > lmax:
            l[int(rng.integers(0, 3))] = 0
        out.append((float(rng.uniform(0.1, 6.0)), l, [float(v) for v in rng.uniform(-1.5, 1.5, 3)]))
    return out


def test_one_electron_primitives_vs_oracle(L, oracle):
    rng = np.random.default_rng(11)
    O = oracle.L
    for _ in range(300):
        (a, la, A), (b, lb, B) = _rand_prims(rng, 2, 3)
        Cc = [float(v) for v in rng.uniform(-1.5, 1.5, 3)]
        args = [a, *la, *A, b, *lb, *B]
        for f in ("overlap", "kinetic"):
            v, w = getattr(L, f)(*args), getattr(O, "or_" + f)(*args)
            assert abs(v - w) <= 1e-12 * max(1.0, abs(w)), (f, args)
        mx = [int(v) for v in rng.integers(0, 3, 3)]
        O.or_moment.argtypes = [D, I, I, I, D, D, D] * 2 + [I, I, I, D, D, D]
        O.or_moment.restype = D
        v, w = L.moment(*args, *mx, *Cc), O.or_moment(*args, *mx, *Cc)
        assert abs(v - w) <= 1e-12 * max(1.0, abs(w)), ("moment", args, mx)
        nargs = [*A, 1.0, *la, a, *B, 1.0, *lb, b, *Cc]
        v, w = L.nai(*nargs), O.or_nai(*nargs)
        assert abs(v - w) <= 1e-11 * max(1.0, abs(w)), nargs


def test_host_eri_vs_oracle(L, oracle):
    rng = np.random.default_rng(12)
    for _ in range(200):
        prims = _rand_prims(rng, 4, 2)
        args = []
        for (e, l, c) in prims:
            args += [*c, 1.0, *l, e]
        v, w = L.eri(*args), oracle.L.or_eri(*args)
        assert abs(v - w) <= 1e-12 * max(1.0, abs(w))


def test_readers_match_oracle(L, oracle):
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    L.readMolecule_XYZ.argtypes = [C.c_void_p]
    L.readMolecule_XYZ.restype = C.POINTER(Molecule)
    L.read_GAMESS_BasisSet.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(I)]
    L.read_GAMESS_BasisSet.restype = C.POINTER(GTOBasisSet)
    L.genBasis.argtypes = [C.POINTER(Molecule), C.POINTER(I), I, C.POINTER(GTOBasisSet)]
    L.genBasis.restype = C.POINTER(GTOBasis)
    xyz = os.path.join(HERE, "golden", "g2", "ch2o_q0m1.xyz").encode()
    f = libc.fopen(xyz, b"r")
    mol = L.readMolecule_XYZ(f)
    libc.fclose(f)
    f = libc.fopen(basis_path("631g_star").encode(), b"r")
    ns = I()
    db = L.read_GAMESS_BasisSet(f, b"631g*", C.byref(ns))
    libc.fclose(f)
    nb = I()
    g = L.genBasis(mol, C.byref(nb), ns.value, db)
    om = oracle.molecule(xyz.decode())
    og, on = oracle.basis(om, basis_path("631g_star"))
    assert nb.value == on == 34  # C, O: 15 each; 2 H: 2 each
    mine, theirs = basis_from_ptr(g, nb.value), basis_from_ptr(og, on)
    for a, b in zip(mine, theirs):
        assert a[:3] == b[:3] and np.allclose(a[3], b[3], rtol=0, atol=1e-14)
        assert np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])
        assert np.allclose(a[6], b[6], rtol=1e-13)
