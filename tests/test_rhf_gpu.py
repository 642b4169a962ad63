"""The legacy RHF surface (src/rhf.h:26-47): rhf() and its two-electron
matrix.

- The G matrix rhf() adds to H, J[P] - K[P]/2 for the RHF density
  P = 2 sum C C^T (built here by sqjk_build on P/2 with the 1e-10 Schwarz
  cut), against the reference's own getGMatrix (src/rhf.c:82-116:
  0.5 * GTO_2JK_Matrix_Symm_Shell, src/matrix.c:854-1127, compiled into
  oracle/_ref).  The reference screens single basis-function pairs at
  1e-10 and this path whole shells, so the two differ by integrals whose
  Schwarz product is below 1e-10: bar 1e-9 max-abs.
- rhf() (damped SCF, stop at |dE| and rms dP <= 1e-6, src/rhf.c:451-510)
  follows the reference's loop restated on the oracle (oracle.legacy_rhf):
  same iteration count, energy to 1e-8 Eh.  That loop stops where the
  energy change per damped step falls below 1e-6, which can be ~2e-6 Eh
  from the converged energy (CH4), so both are also checked against the
  oracle's tightly converged SCF at 1e-5 Eh.
"""
import ctypes as C

import numpy as np
import pytest

from molecules import cluster, g2, random_density
from sqamd import JKEngine, GTOBasis, Molecule, Option, default_option, dptr, lib

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["h2o", "w4"])
def test_legacy_g_matrix_matches_reference(name, reference):
    mol, bas = g2("h2o_q0m1", "631g_star") if name == "h2o" else cluster(4, 401)
    n = bas.n
    P = random_density(n, 7, 0.05)
    sch = reference.schwarz(bas.ptr, n)
    Gr = reference.g_legacy(bas.ptr, n, P, sch)
    eng = JKEngine(bas.ptr, n, schwarz=sch)
    J, K, _ = eng.build(P / 2, P / 2, 0.0, 1e-10, True)
    eng.close()
    d = np.abs(J + K - Gr).max()
    print(f"{name}: n={n} max|G - G_ref| = {d:.2e} (|G| max {np.abs(Gr).max():.3f})")
    assert d < 1e-9


@pytest.mark.parametrize("name", ["h2o_q0m1", "ch4_q0m1"])
def test_legacy_rhf_energy(name, oracle):
    mol, bas = g2(name, "631g_star")
    n, ne = bas.n, mol.n_electrons
    L = lib()
    L.rhf.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int,
                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(Option)]
    L.rhf.restype = C.c_double
    Cm, e = np.zeros((n, n)), np.zeros(n)
    opt = default_option(SCFMax=400)
    E = L.rhf(n, bas.ptr, mol.ptr, ne, dptr(Cm), dptr(e), C.byref(opt))
    from oracle import legacy_rhf
    El, itl = legacy_rhf(oracle, bas.ptr, n, mol.ptr, ne, scf_max=400)
    Eo, _ = oracle.uhf(bas.ptr, n, mol.ptr, ne // 2, ne // 2, restricted=True, scf_conv=1e-10,
                       three_step=False, use_mpole=False, max_mem=0)
    print(f"{name}: rhf() {E:.10f}  restated loop {El:.10f} ({itl} it)  converged SCF {Eo:.10f}")
    assert abs(E - El) < 1e-8
    assert abs(E - Eo) < 1e-5
    assert np.isfinite(e).all() and e[ne // 2 - 1] < e[ne // 2]
