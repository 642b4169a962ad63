"""MP2 on the device (sqjk_eri_pairs, sqjk_mp2_rhf and the
mp2_rhf_aqij_Parallel drop-in) against the oracle (the reference's aqij
algorithm restated in sq_oracle.c, itself pinned against the compiled
reference, the textbook MO-integral sum and the published H2O/STO-3G MP2
energy on the CPU).

Tolerances: integrals of the device store equal the oracle's contracted
integrals to 1e-12 (FP64, different recursions); the unscreened device MP2
energy equals the oracle's exact energy to 1e-10 Eh; at the Schwarz cut-off
1e-10 of the drop-in the device stays within 1e-9 Eh of it, and within
1e-9 Eh of the reference's 1e-8 rules.
"""
import numpy as np
import pytest

from molecules import g2
from sqamd import JKEngine, default_option, inputs
from sqamd.abi import Mol
from sqamd.engine import frozen_core, mp2_rhf_aqij_parallel, mp2_uhf_aqij

pytestmark = pytest.mark.gpu
CRAWFORD_BOHR = np.array([[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
                          [-1.638036840407, 1.136548822547, 0.0]])


def _rhf(oracle, mol, bas, conv=1e-10):
    ne = mol.n_electrons
    _, _, (CA, _, eA, _) = oracle.uhf(bas.ptr, bas.n, mol.ptr, ne // 2, ne // 2, scf_conv=conv, want_orbitals=True)
    return CA, eA


@pytest.mark.parametrize("name,basis", [("h2o_q0m1", "631g_star"), ("nh3_q0m1", "631g_star"), ("h2o_q0m1", "sto3g")])
def test_eri_pairs_vs_oracle(oracle, name, basis):
    """Every contracted integral of the store kernels (all classes of s, p,
    d shells and the L-shell split of STO-3G) against the oracle's."""
    mol, bas = g2(name, basis)
    V = JKEngine(bas.ptr, bas.n).eri_pairs(0.0)
    R = oracle.eri_pairs(bas.ptr, bas.n)
    err = np.abs(V - R).max()
    print(f"{name}/{basis}: m={V.shape[0]} max|dV| = {err:.2e} (max|V| {np.abs(R).max():.3f})")
    assert err < 1e-12
    # a quartet of one pair with itself evaluates (pq|rs) and (rs|pq) as two
    # components, equal up to rounding
    assert np.abs(V - V.T).max() < 1e-13


def test_eri_pairs_dd_classes(oracle):
    """Ethylene 6-31G*: (dd|dd) on the wavefront kernel, primitive-split tiles."""
    mol, bas = g2("c2h4_q0m1", "631g_star")
    V = JKEngine(bas.ptr, bas.n).eri_pairs(0.0)
    R = oracle.eri_pairs(bas.ptr, bas.n)
    err = np.abs(V - R).max()
    print(f"c2h4/6-31G*: max|dV| = {err:.2e}")
    assert err < 1e-12


def test_mp2_water_sto3g_published(oracle):
    """All-electron H2O/STO-3G at the published geometry: -0.049149636120 Eh."""
    mol = Mol([8, 1, 1], CRAWFORD_BOHR)
    bas = inputs.make_basis(mol, "sto3g")
    Cm, e = _rhf(oracle, mol, bas, 1e-12)
    r = JKEngine(bas.ptr, bas.n).mp2_rhf(5, 0, e, Cm, 0.0)
    print(f"E_MP2 = {r['e2']:.12f}")
    assert abs(r["e2"] - (-0.049149636120)) < 1e-11


@pytest.mark.parametrize("name", ["h2o_q0m1", "c2h4_q0m1", "ch3cl_q0m1"])
def test_mp2_exact_vs_oracle(oracle, name):
    mol, bas = g2(name, "631g_star")
    ne = mol.n_electrons
    Cm, e = _rhf(oracle, mol, bas)
    nc = oracle.mp2_frozen_core(mol.ptr)
    r = JKEngine(bas.ptr, bas.n).mp2_rhf(ne // 2, nc, e, Cm, 0.0)
    x = oracle.mp2_rhf(bas.ptr, bas.n, ne // 2, nc, e, Cm, 0.0)
    print(f"{name}: n={bas.n} device {r['e2']:.14f} oracle {x:.14f} d={r['e2'] - x:.1e}")
    assert abs(r["e2"] - x) < 1e-10


def test_mp2_water_dimer_screened(oracle):
    """Two molecules (38 functions): the screened device energy (1e-10)
    against the exact and the reference-rules oracle energies."""
    Z, X = inputs.water_cluster(2, 7)
    mol = inputs.mol_from_angstrom(Z, X)
    bas = inputs.make_basis(mol, "631g_star")
    Cm, e = _rhf(oracle, mol, bas)
    jk = JKEngine(bas.ptr, bas.n)
    ex = jk.mp2_rhf(10, 2, e, Cm, 0.0)["e2"]
    sc = jk.mp2_rhf(10, 2, e, Cm, 1e-10)["e2"]
    x = oracle.mp2_rhf(bas.ptr, bas.n, 10, 2, e, Cm, 0.0)
    rules = oracle.mp2_rhf(bas.ptr, bas.n, 10, 2, e, Cm, 1e-8)
    print(f"exact {ex:.14f} screened {sc:.14f} oracle {x:.14f} reference rules {rules:.14f}")
    assert abs(ex - x) < 1e-10
    assert abs(sc - x) < 1e-9
    assert abs(sc - rules) < 1e-9


def test_mp2_no_correlated_orbitals():
    mol, bas = g2("h2o_q0m1", "sto3g")
    n = bas.n
    r = JKEngine(bas.ptr, n).mp2_rhf(5, 5, np.zeros(n), np.eye(n), 0.0)
    assert r["e2"] == 0.0


def test_mp2_dropin_vs_oracle(oracle, capfd):
    """mp2_rhf_aqij_Parallel: frozen core, the reference's report lines, the
    energy of the reference's rules within 1e-9 Eh."""
    mol, bas = g2("c2h4_q0m1", "631g_star")
    ne = mol.n_electrons
    Cm, e = _rhf(oracle, mol, bas)
    assert frozen_core(mol) == oracle.mp2_frozen_core(mol.ptr) == 2
    e2 = mp2_rhf_aqij_parallel(bas, mol, ne // 2, e, Cm, default_option(restricted=1))
    out = capfd.readouterr().out
    rules = oracle.mp2_rhf(bas.ptr, bas.n, ne // 2, 2, e, Cm, 1e-8)
    print(out)
    print(f"drop-in {e2:.14f} reference rules {rules:.14f}")
    assert "MP2 ENERGY CORRECTION" in out
    assert "The number of excluded core orbitals:          %4d" % 2 in out
    assert f"MP2 Energy Correction is {rules:12.6f} Hartree" in out
    assert abs(e2 - rules) < 1e-9


def _cation(oracle, name, basis="631g_star"):
    mol, bas = g2(name, basis, 1)
    ne = mol.n_electrons
    nea, neb = (ne + 1) // 2, ne // 2
    _, _, (CA, CB, eA, eB) = oracle.uhf(bas.ptr, bas.n, mol.ptr, nea, neb, restricted=False, scf_conv=1e-10,
                                        want_orbitals=True)
    return mol, bas, nea, neb, CA, CB, eA, eB


@pytest.mark.parametrize("name", ["h2o_q0m1", "c2h4_q0m1"])
def test_mp2_uhf_vs_oracle(oracle, name):
    """Open-shell cations (doublets), same-spin and opposite-spin blocks."""
    mol, bas, nea, neb, CA, CB, eA, eB = _cation(oracle, name)
    nc = oracle.mp2_frozen_core(mol.ptr)
    r = JKEngine(bas.ptr, bas.n).mp2_uhf(nea, neb, nc, eA, eB, CA, CB, 0.0)
    x = oracle.mp2_uhf(bas.ptr, bas.n, nea, neb, nc, eA, eB, CA, CB, 0.0)
    print(f"{name}+: device {r['e2']:.14f} oracle {x:.14f} d={r['e2'] - x:.1e}")
    assert abs(r["e2"] - x) < 1e-10


def test_mp2_uhf_closed_shell_equals_rhf(oracle):
    mol, bas = g2("c2h4_q0m1", "631g_star")
    Cm, e = _rhf(oracle, mol, bas)
    jk = JKEngine(bas.ptr, bas.n)
    r = jk.mp2_rhf(8, 2, e, Cm, 0.0)["e2"]
    u = jk.mp2_uhf(8, 8, 2, e, e, Cm, Cm, 0.0)["e2"]
    print(f"RMP2 {r:.15f} UMP2 {u:.15f}")
    assert abs(r - u) < 1e-11


def test_mp2_uhf_dropin_vs_oracle(oracle, capfd):
    mol, bas, nea, neb, CA, CB, eA, eB = _cation(oracle, "h2o_q0m1")
    e2 = mp2_uhf_aqij(bas, mol, nea, neb, eA, eB, CA, CB, default_option(restricted=0))
    out = capfd.readouterr().out
    rules = oracle.mp2_uhf(bas.ptr, bas.n, nea, neb, 1, eA, eB, CA, CB, 1e-8)
    print(out)
    assert "UNRESTRICTED MP2 ENERGY CORRECTION" in out
    assert "The number of included occupied orbitals:  %4d %4d" % (nea - 1, neb - 1) in out
    assert f"MP2 Energy Correction is {rules:12.6f} Hartree" in out
    assert abs(e2 - rules) < 1e-9


def test_mp2_rejects_a_pair_matrix_beyond_device_memory():
    """(H2O)64/6-31G* (1216 functions) would need 4.4 TB for the pair matrix:
    SQJK_ENOMEM with the sizes, before any allocation."""
    Z, X = inputs.water_cluster(64, 11)
    mol = inputs.mol_from_angstrom(Z, X)
    bas = inputs.make_basis(mol, "631g_star")
    n = bas.n
    jk = JKEngine(bas.ptr, n)
    with pytest.raises(RuntimeError, match="MiB of device memory"):
        jk.mp2_rhf(320, 64, np.zeros(n), np.eye(n), 1e-10)


def test_mp2_water16_vs_reference_fixture():
    """(H2O)16/6-31G* (304 functions, 64 correlated and 224 virtual
    orbitals): the device SCF (one-step cut-offs, SCFConv 1e-10) and the
    device MP2 against the reference-J/K SCF orbitals' MP2 energies in
    tests/golden/scf_w16.json: the oracle's exact energy to 1e-9 Eh (the two
    SCFs' orbitals differ at that level), and the reference's own
    mp2_rhf_aqij_Parallel within its screening shift (its 1e-8 rules, which
    the oracle restates to 4e-16, lower the energy by 1.2e-8 Eh here)."""
    import json
    import os
    from molecules import cluster
    from test_grad_gpu import _scf
    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scf_w16.json")))
    mol, bas = cluster(16, 1601)
    ne = mol.n_electrons
    E, (CA, _, eA, _), _ = _scf(mol, bas, ne // 2, ne // 2, conv=1e-10)
    r = JKEngine(bas.ptr, bas.n).mp2_rhf(ne // 2, fx["mp2_frozen_core"], eA, CA, 1e-10)
    exact, ref = fx["mp2_energy_oracle_exact"], fx["mp2_energy"]
    print(f"E_SCF {E:.10f} (reference {fx['energy']:.10f})  E2 {r['e2']:.12f} (exact {exact:.12f}, "
          f"reference rules {ref:.12f})")
    assert abs(E - fx["energy"]) < 1e-8
    assert abs(r["e2"] - exact) < 1e-9
    assert abs(r["e2"] - ref) < abs(exact - ref) + 1e-9
