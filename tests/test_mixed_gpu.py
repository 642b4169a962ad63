"""Mixed precision (FP32 quartets below a Schwarz threshold, FP64 digestion):
J/K against the FP64 device path and the oracle, and SCF |dE| against the
FP64 SCF (|dE| <= 1e-6 Eh per molecule)."""
import ctypes as C

import numpy as np
import pytest

from molecules import cluster, g2, random_density
from sqamd import GTOBasis, JKEngine, Molecule, Option, default_option, dptr, lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("thr", [1e-6, 1e-4, 1e-2])
def test_mixed_jk_close_to_fp64(oracle, thr):
    mol, bas = cluster(2, 11, "631g_star")
    n = bas.n
    sch = oracle.schwarz(bas.ptr, n)
    P = random_density(n, 4)
    Jo, Ko, _ = oracle.jk(oracle.jk_context(False, 0), bas.ptr, n, P, P, sch, 0.0, 1e-12, True)
    eng = JKEngine(bas.ptr, n, schwarz=sch)
    eng.set_mixed(thr)
    J, K, _ = eng.build(P, P, 0.0, 1e-12, True)
    st = eng.stats()
    eng.close()
    dj, dk = np.abs(J - Jo).max(), np.abs(K - Ko).max()
    print(f"thr={thr:g}: fp32 quartets {st['quartets_fp32']}/{st['quartets']}  dJ={dj:.2e} dK={dk:.2e}")
    assert st["quartets_fp32"] > 0
    # FP32 integrals carry ~1e-7 relative error and are bounded by thr
    assert dj < 50 * thr * 1e-6 + 1e-10 and dk < 50 * thr * 1e-6 + 1e-10


def _uhf(mol, bas, nE, mixed):
    L = lib()
    L.sqjk_set_dropin_mixed.argtypes = [C.c_double]
    L.uhf.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int, C.c_int] + \
        [C.POINTER(C.c_double)] * 4 + [C.POINTER(Option)]
    L.uhf.restype = C.c_double
    n = bas.n
    CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
    opt = default_option(SCFConv=1e-9, SCFAccuracy=1)
    L.sqjk_set_dropin_mixed(mixed)
    try:
        return L.uhf(n, bas.ptr, mol.ptr, nE, nE, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt))
    finally:
        L.sqjk_set_dropin_mixed(0.0)


@pytest.mark.parametrize("name", ["h2o_q0m1", "c2h4_q0m1"])
def test_mixed_scf_energy_within_1uEh(name):
    mol, bas = g2(name, "631g_star")
    ne = mol.n_electrons // 2
    E64 = _uhf(mol, bas, ne, 0.0)
    Emx = _uhf(mol, bas, ne, 1e-2)
    print(f"{name}: E(fp64) = {E64:.10f}  E(mixed 1e-2) = {Emx:.10f}  |dE| = {abs(Emx - E64):.2e}")
    assert E64 != 0.0 and abs(Emx - E64) <= 1e-6
