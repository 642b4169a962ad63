"""Edge cases and full-size checks of the HIP path.

- Textbook STO-3G energies through the uhf() drop-in: H atom (one basis
  function, no beta electron), He atom, H2 at R = 1.4 bohr (Szabo and
  Ostlund, Modern Quantum Chemistry, ch. 3: -1.1167 Eh; H -0.46658 Eh,
  He -2.80778 Eh), and the oracle's SCF on the same inputs.
- J/K of a single basis function and of (H2O)8/6-31G* against the oracle.
- (H2O)32/6-31G* (608 functions, the benchmark size): properties that do not
  need the oracle: symmetry, sign of the Coulomb and exchange energies,
  linearity in the density with density-independent screening, and the UHF
  build with equal spin densities equal to the RHF build.
"""
import ctypes as C

import numpy as np
import pytest

from molecules import cluster, random_density
from sqamd import GTOBasis, JKEngine, Mol, Molecule, Option, default_option, dptr, inputs, lib

pytestmark = pytest.mark.gpu


def _uhf(mol, bas, nEA, nEB, **kw):
    L = lib()
    L.uhf.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int, C.c_int] + \
        [C.POINTER(C.c_double)] * 4 + [C.POINTER(Option)]
    L.uhf.restype = C.c_double
    n = bas.n
    CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
    opt = default_option(**kw)
    return L.uhf(n, bas.ptr, mol.ptr, nEA, nEB, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt))


@pytest.mark.parametrize("name,Z,X,nea,neb,textbook", [
    ("H", [1], [[0, 0, 0]], 1, 0, -0.46658185),
    ("He", [2], [[0, 0, 0]], 1, 1, -2.80778396),
    ("H2", [1, 1], [[0, 0, 0], [0, 0, 1.4]], 1, 1, -1.11671433),
])
def test_textbook_sto3g(oracle, name, Z, X, nea, neb, textbook):
    mol = Mol(Z, X)
    bas = inputs.make_basis(mol, "sto3g")
    restricted = int(nea == neb)
    E = _uhf(mol, bas, nea, neb, SCFConv=1e-10, SCFAccuracy=1, restricted=restricted)
    Eo, _ = oracle.uhf(bas.ptr, bas.n, mol.ptr, nea, neb, restricted=bool(restricted), scf_conv=1e-10,
                       three_step=False, use_mpole=False, max_mem=0)
    print(f"{name}: n={bas.n} E(gpu)={E:.10f} E(oracle)={Eo:.10f} textbook={textbook}")
    assert abs(E - Eo) < 1e-9
    assert abs(E - textbook) < 1e-7


def _vs_oracle(oracle, bas, P, cut, tol):
    n = bas.n
    sch = oracle.schwarz(bas.ptr, n)
    Jo, Ko, _ = oracle.jk(oracle.jk_context(False, 0), bas.ptr, n, P, P, sch, cut, cut, True)
    eng = JKEngine(bas.ptr, n, schwarz=sch)
    J, K, _ = eng.build(P, P, cut, cut, True)
    st = eng.stats()
    eng.close()
    dJ, dK = np.abs(J - Jo).max(), np.abs(K - Ko).max()
    print(f"n={n} cut={cut:g} quartets={st['quartets']} dJ={dJ:.2e} dK={dK:.2e}")
    assert dJ < tol and dK < tol


def test_single_function(oracle):
    mol = Mol([1], [[0.1, -0.2, 0.3]])
    bas = inputs.make_basis(mol, "sto3g")
    _vs_oracle(oracle, bas, np.array([[0.7]]), 0.0, 1e-13)


def test_water8_screened(oracle):
    """(H2O)8/6-31G* (152 functions), the benchmark's screening rule at 1e-12."""
    mol, bas = cluster(8, 801)
    _vs_oracle(oracle, bas, random_density(bas.n, 8, scale=0.01), 1e-12, 1e-10)


@pytest.fixture(scope="module")
def big():
    mol, bas = cluster(32, 3201)
    eng = JKEngine(bas.ptr, bas.n)
    yield bas, eng
    eng.close()


def test_full_size_symmetry_and_signs(big):
    bas, eng = big
    n = bas.n
    P = random_density(n, 32, scale=0.01)
    P = P @ P.T / n  # positive semidefinite, like a density
    J, K, _ = eng.build(P, P, 1e-12, 1e-12, True)
    assert np.array_equal(J, J.T) and np.array_equal(K, K.T)
    eJ, eK = np.sum(P * J), np.sum(P * K)
    print(f"n={n} Tr(PJ)={eJ:.6e} Tr(PK)={eK:.6e} quartets={eng.stats()['quartets']}")
    assert eJ > 0.0 and eK < 0.0


def test_full_size_linear_in_density(big):
    """With Schwarz-only screening (fixedCutoff 0) the screened integral set
    does not depend on the density, so J and K are linear in it."""
    bas, eng = big
    n = bas.n
    P1, P2 = random_density(n, 1, scale=0.01), random_density(n, 2, scale=0.01)
    J1, K1, _ = eng.build(P1, P1, 0.0, 1e-12, True)
    J2, K2, _ = eng.build(P2, P2, 0.0, 1e-12, True)
    J3, K3, _ = eng.build(P1 + 2 * P2, P1 + 2 * P2, 0.0, 1e-12, True)
    scale = max(np.abs(J3).max(), 1.0)
    dJ = np.abs(J3 - (J1 + 2 * J2)).max() / scale
    dK = np.abs(K3 - (K1 + 2 * K2)).max() / max(np.abs(K3).max(), 1.0)
    print(f"linearity: rel dJ={dJ:.2e} rel dK={dK:.2e}")
    assert dJ < 1e-12 and dK < 1e-12


def test_full_size_uhf_equals_rhf(big):
    bas, eng = big
    n = bas.n
    P = random_density(n, 9, scale=0.01)
    J, K, _ = eng.build(P, P, 1e-12, 1e-12, True)
    Ju, KA, KB = eng.build(P, P, 1e-12, 1e-12, False)
    d = max(np.abs(Ju - J).max(), np.abs(KA - K).max(), np.abs(KB - K).max())
    print(f"UHF(PA=PB) vs RHF: max diff {d:.2e}")
    assert d < 1e-11
