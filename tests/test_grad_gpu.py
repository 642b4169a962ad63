"""Energy gradient on the device (sqjk_grad_eri, sqjk_grad_one_electron and
the uhf_force / GradEri_ShellSet_Parallel drop-ins) against the oracle
(the CPU restatement of src/grad.c, itself pinned against the compiled
reference and finite differences on the CPU) and against finite differences
of the device SCF energy.

Tolerances: the device two-electron gradient without screening equals the
oracle's exact derivative to 1e-10 (FP64, different summation order); with
the reference's 1e-10 cut-off the device stays within 1e-9 of the exact
gradient, while the reference's rules (which also drop single primitive
terms below the cut) move it by up to 2e-7, the bound used against them;
forces from finite differences of SCF energies converged to 1e-11,
2e-6 Eh/bohr.
"""
import ctypes as C

import numpy as np
import pytest

from molecules import g2, random_density, water
from sqamd import JKEngine, default_option, inputs
from sqamd.engine import basis_to_atom, grad_one_electron, uhf_force

pytestmark = pytest.mark.gpu
BOHR = 1.8897261245650618


def _densities(n, seed, uhf=True):
    PA = random_density(n, seed, 0.2)
    PB = random_density(n, seed + 1, 0.2) if uhf else PA
    return PA, PB


@pytest.mark.parametrize("name,basis", [("h2o_q0m1", "631g_star"), ("c2h4_q0m1", "631g_star"),
                                        ("h2o_q0m1", "sto3g")])
@pytest.mark.parametrize("kfactor", [1.0, 0.5])
def test_grad_eri_exact_vs_oracle(oracle, name, basis, kfactor):
    """Unscreened device gradient = the oracle's exact derivative (every
    class up to (dd|dd) for 6-31G*, UHF densities)."""
    mol, bas = g2(name, basis)
    n, na = bas.n, mol.ptr.contents.nAtom
    b2a = basis_to_atom(mol, bas)
    PA, PB = _densities(n, 11)
    jk = JKEngine(bas.ptr, n)
    G = jk.grad_eri(PA, PB, b2a, na, kfactor, cutoff=0.0)
    R = oracle.grad_eri(bas.ptr, n, b2a, na, PA + PB, PA - PB, kfactor, cutoff=0.0)
    err = np.abs(G - R).max()
    print(f"{name}/{basis} k={kfactor}: n={n} max|dG| = {err:.2e} (max|G| {np.abs(R).max():.3f})")
    assert err < 1e-10 * max(1.0, np.abs(R).max())
    assert np.abs(G.sum(0)).max() < 1e-10


def test_grad_eri_screened_vs_reference_rules(oracle):
    """At the reference's cut-off (1e-10).  The reference's rules (restated
    by the oracle) also drop single primitive terms below the cut, which
    moves this gradient by 2e-7 from the exact one; the device drops only
    whole quartets below it, so it stays within 1e-9 of the exact gradient
    and within the reference's own deviation of the reference."""
    mol, bas = g2("c2h4_q0m1", "631g_star")
    n, na = bas.n, mol.ptr.contents.nAtom
    b2a = basis_to_atom(mol, bas)
    PA, PB = _densities(n, 21, uhf=False)
    jk = JKEngine(bas.ptr, n)
    G = jk.grad_eri(PA, PB, b2a, na, 1.0, cutoff=1e-10)
    R = oracle.grad_eri(bas.ptr, n, b2a, na, PA + PB, PA - PB, 1.0)
    X = oracle.grad_eri(bas.ptr, n, b2a, na, PA + PB, PA - PB, 1.0, cutoff=0.0)
    ref_dev = np.abs(R - X).max()
    print(f"max|G - exact| = {np.abs(G - X).max():.2e}  max|G - ref| = {np.abs(G - R).max():.2e}  "
          f"reference rules vs exact {ref_dev:.2e}")
    assert np.abs(G - X).max() < 1e-9
    assert np.abs(G - R).max() < ref_dev + 1e-9


def test_grad_eri_water_dimer(oracle):
    """Two molecules, 38 functions: pairs across centres and the tile splits
    of longer pair lists."""
    Z, X = inputs.water_cluster(2, 7)
    mol = inputs.mol_from_angstrom(Z, X)
    bas = inputs.make_basis(mol, "631g_star")
    n, na = bas.n, len(Z)
    b2a = basis_to_atom(mol, bas)
    PA, PB = _densities(n, 31)
    G = JKEngine(bas.ptr, n).grad_eri(PA, PB, b2a, na, 1.0, cutoff=0.0)
    R = oracle.grad_eri(bas.ptr, n, b2a, na, PA + PB, PA - PB, 1.0, cutoff=0.0)
    print(f"max|dG| = {np.abs(G - R).max():.2e}")
    assert np.abs(G - R).max() < 1e-10 * max(1.0, np.abs(R).max())


@pytest.mark.parametrize("name,basis", [("h2o_q0m1", "631g_star"), ("c2h4_q0m1", "631g_star")])
def test_grad_one_electron_vs_oracle(oracle, name, basis):
    mol, bas = g2(name, basis)
    n = bas.n
    b2a = basis_to_atom(mol, bas)
    PT, QT = random_density(n, 41, 0.3), random_density(n, 42, 0.3)
    G = grad_one_electron(bas, mol, PT, QT, b2a)
    R = oracle.grad_one_electron(bas.ptr, n, mol.ptr, b2a, PT, QT)
    err = np.abs(G - R).max()
    print(f"{name}/{basis}: max|dG| = {err:.2e} (max|G| {np.abs(R).max():.3f})")
    assert err < 1e-10 * max(1.0, np.abs(R).max())


def _scf(mol, bas, nEA, nEB, conv=1e-11):
    from test_scf_gpu import _uhf  # noqa: F401  (binding of the uhf drop-in)
    from sqamd import GTOBasis, Molecule, Option, dptr, lib
    L = lib()
    L.uhf.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int, C.c_int] + \
        [C.POINTER(C.c_double)] * 4 + [C.POINTER(Option)]
    L.uhf.restype = C.c_double
    n = bas.n
    CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
    opt = default_option(SCFConv=conv, SCFAccuracy=1, restricted=int(nEA == nEB))
    E = L.uhf(n, bas.ptr, mol.ptr, nEA, nEB, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt))
    return E, (CA, CB, eA, eB), opt


@pytest.mark.parametrize("basis,charge,mult", [("sto3g", 0, 1), ("631g_star", 0, 1), ("631g_star", 1, 2)])
def test_uhf_force_vs_oracle(oracle, basis, charge, mult):
    """The uhf_force drop-in on the device SCF's orbitals = the oracle's
    restatement of src/grad.c:2155-2420 on the same orbitals."""
    Z, X = inputs.water_c1()
    mol = inputs.mol_from_angstrom(Z, X, charge)
    bas = inputs.make_basis(mol, basis)
    ne = int(sum(Z)) - charge
    nEA = (ne + mult - 1) // 2
    E, orbs, opt = _scf(mol, bas, nEA, ne - nEA)
    F = uhf_force(bas, mol, nEA, ne - nEA, *orbs, opt)
    R = oracle.force(bas.ptr, bas.n, mol.ptr, nEA, ne - nEA, orbs)
    print(f"{basis} q={charge}: E={E:.10f} max|dF| = {np.abs(F - R).max():.2e}")
    assert np.abs(F - R).max() < 1e-7
    assert np.abs(F.sum(0)).max() < 1e-8


def test_uhf_force_vs_finite_differences():
    """Forces of the device SCF = -dE/dR of the device SCF energy (central
    differences, h = 1e-4 bohr, three coordinates of water/6-31G*)."""
    Z, X = inputs.water_c1()
    mol = inputs.mol_from_angstrom(Z, X)
    bas = inputs.make_basis(mol, "631g_star")
    E, orbs, opt = _scf(mol, bas, 5, 5)
    F = uhf_force(bas, mol, 5, 5, *orbs, opt)
    h = 1e-4 / BOHR
    for a, d in ((0, 1), (1, 0), (2, 1)):
        Xp, Xm = X.copy(), X.copy()
        Xp[a, d] += h
        Xm[a, d] -= h
        Ep = _scf(inputs.mol_from_angstrom(Z, Xp), inputs.make_basis(inputs.mol_from_angstrom(Z, Xp), "631g_star"),
                  5, 5)[0]
        Em = _scf(inputs.mol_from_angstrom(Z, Xm), inputs.make_basis(inputs.mol_from_angstrom(Z, Xm), "631g_star"),
                  5, 5)[0]
        fd = -(Ep - Em) / 2e-4
        print(f"atom {a} axis {d}: F = {F[a, d]:.8f}  -dE/dR = {fd:.8f}")
        assert abs(fd - F[a, d]) < 2e-6


def test_grad_eri_dropin_children_sum(oracle):
    """GradEri_ShellSet_Parallel: childID 0 adds everything, other children
    add nothing; the sum over children equals the oracle's."""
    from sqamd import GTOBasis, dptr, lib
    mol, bas = g2("h2o_q0m1", "631g_star")
    n, na = bas.n, mol.ptr.contents.nAtom
    b2a = basis_to_atom(mol, bas)
    PA, PB = _densities(n, 51)
    PT, PS = PA + PB, PA - PB
    L = lib()
    PD = C.POINTER(C.c_double)
    L.GradEri_ShellSet_Parallel.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(GTOBasis), C.POINTER(C.c_int),
                                            PD, PD, C.c_double, PD, PD, PD]
    G = np.zeros((3, na))
    for child in range(3):
        L.GradEri_ShellSet_Parallel(child, 3, n, bas.ptr, b2a.ctypes.data_as(C.POINTER(C.c_int)), dptr(PT),
                                    dptr(PS), 1.0, dptr(G[0]), dptr(G[1]), dptr(G[2]))
    R = oracle.grad_eri(bas.ptr, n, b2a, na, PT, PS, 1.0)
    assert np.abs(G.T - R).max() < 1e-7
