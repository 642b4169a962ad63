"""Device one-electron matrices (sqjk_one_electron) against the host
integrals GTO_kinetic, GTO_nuclei and GTO_overlap (pinned against the oracle
in test_host_c.py), element by element, for every pair class up to (d,d)."""
import ctypes as C

import numpy as np
import pytest

from molecules import cluster, g2
from sqamd import GTOBasis, Molecule, lib

pytestmark = pytest.mark.gpu


def _host(bas, mol):
    L = lib()
    for fn in ("GTO_overlap", "GTO_kinetic"):
        getattr(L, fn).argtypes = [C.c_int, C.c_int, C.POINTER(GTOBasis)]
        getattr(L, fn).restype = C.c_double
    L.GTO_nuclei.argtypes = [C.c_int, C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule)]
    L.GTO_nuclei.restype = C.c_double
    n = bas.n
    T, V, S = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            T[i, j] = T[j, i] = L.GTO_kinetic(i, j, bas.ptr)
            V[i, j] = V[j, i] = L.GTO_nuclei(i, j, bas.ptr, mol.ptr)
            S[i, j] = S[j, i] = L.GTO_overlap(i, j, bas.ptr)
    return T, V, S


def _device(bas, mol):
    L = lib()
    PD = C.POINTER(C.c_double)
    L.sqjk_one_electron.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int, PD, PD, PD]
    n = bas.n
    T, V, S = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    rc = L.sqjk_one_electron(n, bas.ptr, mol.ptr, 0, T.ctypes.data_as(PD), V.ctypes.data_as(PD),
                             S.ctypes.data_as(PD))
    assert rc == 0
    return T, V, S


@pytest.mark.parametrize("name,basis", [("c2h4_q0m1", "631g_star"), ("h2o_q0m1", "sto3g"),
                                        ("ch2o_q0m1", "631g_star")])
def test_one_electron_matches_host(name, basis):
    mol, bas = g2(name, basis)
    Th, Vh, Sh = _host(bas, mol)
    Td, Vd, Sd = _device(bas, mol)
    d = [np.abs(Td - Th).max(), np.abs(Vd - Vh).max(), np.abs(Sd - Sh).max()]
    print(f"{name}/{basis} n={bas.n}: dT={d[0]:.2e} dV={d[1]:.2e} dS={d[2]:.2e}")
    assert np.array_equal(Td, Td.T) and np.array_equal(Vd, Vd.T) and np.array_equal(Sd, Sd.T)
    assert max(d) < 1e-12


def test_one_electron_cluster():
    """(H2O)4: far atom-pair combinations (large Boys arguments)."""
    mol, bas = cluster(4, 401)
    Th, Vh, Sh = _host(bas, mol)
    Td, Vd, Sd = _device(bas, mol)
    d = [np.abs(Td - Th).max(), np.abs(Vd - Vh).max(), np.abs(Sd - Sh).max()]
    print(f"(H2O)4 n={bas.n}: dT={d[0]:.2e} dV={d[1]:.2e} dS={d[2]:.2e}")
    assert max(d) < 1e-11
