"""The device-resident SCF (sqscf_*): convergence methods, geometry steps.

- Every opt->convMethod of the reference (damping, 2-, 3-, 4-point density
  DIIS, src/conv.c) converges to the same energy: the SCF fixed point does
  not depend on the path.  Bar 1e-8 Eh at SCFConv 1e-10.
- A geometry step on the same context (sqjk_update_geometry keeps the J/K
  context; the previous orbitals, S-normalised in the new S, give the guess
  as in src/qmd.c:585-589) reaches the energy of a fresh solve at that
  geometry, in fewer iterations.
- A J/K context moved to a new geometry builds the same J/K as a context
  created there.
"""
import ctypes as C

import numpy as np
import pytest

from molecules import g2, random_density
from sqamd import JKEngine, GTOBasis, Molecule, Option, default_option, dptr, inputs, lib
from sqamd.abi import CONVMETHOD_DAMPING, CONVMETHOD_DIIS2, CONVMETHOD_DIIS3, CONVMETHOD_DIIS4
from sqamd.scf import DeviceSCF

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


@pytest.mark.parametrize("method,maxit", [(CONVMETHOD_DAMPING, 400), (CONVMETHOD_DIIS2, 120),
                                          (CONVMETHOD_DIIS3, 120)])
def test_convergence_methods_same_fixed_point(method, maxit):
    mol, bas = g2("h2o_q0m1", "631g_star")
    E4 = _uhf(mol, bas, 5, 5, SCFConv=1e-10, SCFAccuracy=1, convMethod=CONVMETHOD_DIIS4)
    Em = _uhf(mol, bas, 5, 5, SCFConv=1e-10, SCFAccuracy=1, convMethod=method, SCFMax=maxit)
    print(f"method {method}: E = {Em:.10f}  DIIS4 E = {E4:.10f}")
    assert E4 != 0.0 and Em != 0.0
    assert abs(Em - E4) < 1e-8


def test_uhf_doublet_methods():
    mol, bas = g2("ch3_q0m2", "631g_star")
    E4 = _uhf(mol, bas, 5, 4, SCFConv=1e-10, SCFAccuracy=1, restricted=0)
    E3 = _uhf(mol, bas, 5, 4, SCFConv=1e-10, SCFAccuracy=1, restricted=0, convMethod=CONVMETHOD_DIIS3,
              SCFMax=150)
    print(f"CH3 UHF: DIIS4 {E4:.10f} DIIS3 {E3:.10f}")
    assert E4 != 0.0 and abs(E3 - E4) < 1e-8


def _displaced(mol0, seed, sigma_ang=0.05):
    rng = np.random.default_rng(seed)
    xyz = np.stack([mol0.x, mol0.y, mol0.z], axis=1)
    xyz = xyz + rng.normal(0.0, sigma_ang * inputs.ANGSTROM2BOHR, xyz.shape)
    from sqamd import Mol
    return Mol(mol0.Z, xyz)


def test_geometry_steps_reuse_context():
    mol0, bas0 = g2("ch2o_q0m1", "631g_star")
    ne = mol0.n_electrons
    scf = DeviceSCF()
    H, S = scf.set_system(bas0, mol0)
    scf.diag_guess(S, ne // 2, ne // 2)
    r0 = scf.run(ne // 2, ne // 2, conv=1e-10, three_step=False)
    E0 = _uhf(mol0, bas0, ne // 2, ne // 2, SCFConv=1e-10, SCFAccuracy=1)
    print(f"step 0: E {r0['energy']:.10f} ({r0['iterations']} it)  uhf() {E0:.10f}")
    assert abs(r0["energy"] - E0) < 1e-8
    for step in (1, 2):
        mol = _displaced(mol0, 100 + step, sigma_ang=0.02)
        bas = inputs.make_basis(mol, "631g_star")
        scf.set_system(bas, mol)
        scf.guess_from_orbitals(ne // 2, ne // 2)
        r = scf.run(ne // 2, ne // 2, conv=1e-10, three_step=False)
        Ef = _uhf(mol, bas, ne // 2, ne // 2, SCFConv=1e-10, SCFAccuracy=1)
        # the same geometry from the DIAG guess on a fresh context
        fresh = DeviceSCF()
        _, S1 = fresh.set_system(bas, mol)
        fresh.diag_guess(S1, ne // 2, ne // 2)
        rf = fresh.run(ne // 2, ne // 2, conv=1e-10, three_step=False)
        fresh.close()
        print(f"step {step}: E {r['energy']:.10f} ({r['iterations']} it)  fresh uhf() {Ef:.10f}  "
              f"fresh DIAG start {rf['energy']:.10f} ({rf['iterations']} it)")
        assert r["converged"] and abs(r["energy"] - Ef) < 1e-8 and abs(rf["energy"] - Ef) < 1e-8
        assert r["iterations"] < rf["iterations"]
    scf.close()


def test_update_geometry_matches_fresh_context():
    mol0, bas0 = g2("c2h4_q0m1", "631g_star")
    mol1 = _displaced(mol0, 7)
    bas1 = inputs.make_basis(mol1, "631g_star")
    n = bas0.n
    P = random_density(n, 11)
    eng = JKEngine(bas0.ptr, n)
    eng.build(P, P, 1e-11, 1e-11, True)
    L = lib()
    L.sqjk_update_geometry.argtypes = [C.c_void_p, C.c_int, C.POINTER(GTOBasis)]
    assert L.sqjk_update_geometry(eng._h, bas1.n, bas1.ptr) == 0
    J1, K1, _ = eng.build(P, P, 1e-11, 1e-11, True)
    s1 = eng.schwarz()
    eng.close()
    fresh = JKEngine(bas1.ptr, n)
    J2, K2, _ = fresh.build(P, P, 1e-11, 1e-11, True)
    s2 = fresh.schwarz()
    fresh.close()
    d = max(np.abs(J1 - J2).max(), np.abs(K1 - K2).max(), np.abs(s1 - s2).max())
    print(f"moved vs fresh context: max diff {d:.2e}")
    assert d < 1e-12
    # a different basis is refused: another size, or the same functions
    # with one exponent changed
    mol2, bas2 = g2("c2h4_q0m1", "sto3g")
    eng = JKEngine(bas0.ptr, n)
    assert L.sqjk_update_geometry(eng._h, bas2.n, bas2.ptr) != 0
    bas3 = inputs.make_basis(mol1, "631g_star")
    bas3.arr[0].exp[0] *= 1.01
    assert L.sqjk_update_geometry(eng._h, bas3.n, bas3.ptr) != 0
    eng.close()


def test_context_switches_basis():
    """One SCF context serving molecules of different bases in turn (as a
    serving slot does): each solve equals uhf() on its own."""
    scf = DeviceSCF()
    for name, basis in (("h2o_q0m1", "631g_star"), ("c2h4_q0m1", "631g_star"), ("h2o_q0m1", "sto3g"),
                        ("ch2o_q0m1", "631g_star")):
        mol, bas = g2(name, basis)
        ne = mol.n_electrons
        _, S = scf.set_system(bas, mol)
        scf.diag_guess(S, ne // 2, ne // 2)
        r = scf.run(ne // 2, ne // 2, conv=1e-10, three_step=False)
        E = _uhf(mol, bas, ne // 2, ne // 2, SCFConv=1e-10, SCFAccuracy=1)
        print(f"{name}/{basis} n={bas.n}: {r['energy']:.10f} ({r['iterations']} it), uhf() {E:.10f}")
        assert r["converged"] and abs(r["energy"] - E) < 1e-8
    scf.close()
