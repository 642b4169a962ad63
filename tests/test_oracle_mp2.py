"""The oracle's MP2 (sq_oracle.c, the reference's aqij algorithm of
src/mp2.c:581-712 and src/matrix.c:4797-5073) pinned three ways:

- the published all-electron MP2 energy of H2O/STO-3G at the geometry of
  T. D. Crawford's programming projects (project #4: E_SCF =
  -74.942079928192, E_MP2 = -0.049149636120 Eh);
- the textbook sum over MO integrals from the full AO tensor (einsum
  transform, Szabo-Ostlund eq. 6.74), exact and with a frozen core;
- the compiled reference's own mp2_rhf_aqij_Parallel (oracle/_ref) at its
  MP2_SCHWARZ_CUTOFF rules (1e-8), water and ethylene in 6-31G*.

The unrestricted MP2 (src/mp2.c:2126-2432) is pinned the same way: against
mp2_uhf_aqij of the compiled reference (the water cation), the spin-orbital
textbook sum, and the closed-shell identity E_UMP2(C, C) = E_RMP2(C).
"""
import os

import numpy as np
import pytest

from molecules import g2
from sqamd import inputs
from sqamd.abi import Mol

HERE = os.path.dirname(os.path.abspath(__file__))
CRAWFORD_BOHR = np.array([[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
                          [-1.638036840407, 1.136548822547, 0.0]])


def crawford_water():
    mol = Mol([8, 1, 1], CRAWFORD_BOHR)
    return mol, inputs.make_basis(mol, "sto3g")


def _rhf(oracle, mol, bas, conv=1e-12):
    ne = mol.n_electrons
    E, _, (CA, _, eA, _) = oracle.uhf(bas.ptr, bas.n, mol.ptr, ne // 2, ne // 2, scf_conv=conv, want_orbitals=True)
    return E, CA, eA


def mp2_textbook(oracle, bas, Cm, e, nocc, ncore):
    """E2 = sum_abrs (ar|bs)[2(ar|bs) - (as|br)]/(e_a+e_b-e_r-e_s) from the
    full AO tensor, one quarter transform at a time."""
    n = bas.n
    V = oracle.eri_pairs(bas.ptr, n)
    idx = np.zeros((n, n), dtype=int)
    for p in range(n):
        for q in range(p + 1):
            idx[p, q] = idx[q, p] = p * (p + 1) // 2 + q
    T = V[idx[:, :, None, None], idx[None, None, :, :]]
    Co, Cv = Cm.T[:, ncore:nocc], Cm.T[:, nocc:]
    g = np.einsum("pqrs,pa->aqrs", T, Co)
    g = np.einsum("aqrs,qi->airs", g, Cv)
    g = np.einsum("airs,rb->aibs", g, Co)
    g = np.einsum("aibs,sj->aibj", g, Cv)
    eo, ev = e[ncore:nocc], e[nocc:]
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return float(np.sum(g * (2.0 * g - g.transpose(0, 3, 2, 1)) / D))


def test_mp2_water_sto3g_published(oracle):
    mol, bas = crawford_water()
    E, Cm, e = _rhf(oracle, mol, bas)
    assert abs(E - (-74.942079928192)) < 1e-9
    e2 = oracle.mp2_rhf(bas.ptr, bas.n, 5, 0, e, Cm, 0.0)
    print(f"E_SCF {E:.12f}  E_MP2 {e2:.12f}")
    assert abs(e2 - (-0.049149636120)) < 1e-11


@pytest.mark.parametrize("ncore", [0, 1])
def test_mp2_oracle_vs_textbook_sum(oracle, ncore):
    mol, bas = g2("h2o_q0m1", "631g_star")
    _, Cm, e = _rhf(oracle, mol, bas, 1e-10)
    a = oracle.mp2_rhf(bas.ptr, bas.n, 5, ncore, e, Cm, 0.0)
    b = mp2_textbook(oracle, bas, Cm, e, 5, ncore)
    print(f"ncore {ncore}: oracle {a:.15f} textbook {b:.15f}")
    assert abs(a - b) < 1e-13


def test_mp2_published_value_catches_a_dropped_exchange_transpose(oracle):
    """A planted mistake in the textbook sum (the exchange integral (as|br)
    read as (ar|bs)) misses the published energy by far more than the
    tolerance, so the published pin would catch it."""
    mol, bas = crawford_water()
    _, Cm, e = _rhf(oracle, mol, bas)
    n = bas.n
    V = oracle.eri_pairs(bas.ptr, n)
    idx = np.array([[max(p, q) * (max(p, q) + 1) // 2 + min(p, q) for q in range(n)] for p in range(n)])
    T = V[idx[:, :, None, None], idx[None, None, :, :]]
    Co, Cv = Cm.T[:, :5], Cm.T[:, 5:]
    g = np.einsum("pqrs,pa,qi,rb,sj->aibj", T, Co, Cv, Co, Cv)
    eo, ev = e[:5], e[5:]
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    right = float(np.sum(g * (2.0 * g - g.transpose(0, 3, 2, 1)) / D))
    wrong = float(np.sum(g * (2.0 * g - g) / D))
    assert abs(right - (-0.049149636120)) < 1e-11
    assert abs(wrong - (-0.049149636120)) > 1e-3


def test_frozen_core_counts(oracle, reference_mp2):
    for name in ("h2o_q0m1", "c2h4_q0m1", "hcl_q0m1", "sih4_q0m1"):
        path = os.path.join(HERE, "golden", "g2", name + ".xyz")
        if not os.path.exists(path):
            continue
        mol, _ = g2(name, "sto3g")
        assert oracle.mp2_frozen_core(mol.ptr) == reference_mp2.L.getFrozenCore_RHF(mol.ptr)


@pytest.mark.parametrize("name", ["h2o_q0m1", "c2h4_q0m1"])
def test_mp2_oracle_vs_reference(oracle, reference_mp2, name):
    """At the reference's rules (cut 1e-8) the oracle equals
    mp2_rhf_aqij_Parallel; without them it moves by at most 1e-11 here."""
    mol, bas = g2(name, "631g_star")
    ne = mol.n_electrons
    _, Cm, e = _rhf(oracle, mol, bas, 1e-10)
    nc = oracle.mp2_frozen_core(mol.ptr)
    rules = oracle.mp2_rhf(bas.ptr, bas.n, ne // 2, nc, e, Cm, 1e-8)
    exact = oracle.mp2_rhf(bas.ptr, bas.n, ne // 2, nc, e, Cm, 0.0)
    ref = reference_mp2.mp2(bas.ptr, bas.n, mol.ptr, ne // 2, e, Cm)
    print(f"{name}: rules {rules:.15f} reference {ref:.15f} exact {exact:.15f}")
    assert abs(rules - ref) < 1e-14
    assert abs(exact - ref) < 1e-11


def _water_cation(oracle):
    Z, X = inputs.water_c1()
    mol = inputs.mol_from_angstrom(Z, X, 1)
    bas = inputs.make_basis(mol, "631g_star")
    _, _, (CA, CB, eA, eB) = oracle.uhf(bas.ptr, bas.n, mol.ptr, 5, 4, restricted=False, scf_conv=1e-10,
                                        want_orbitals=True)
    return mol, bas, CA, CB, eA, eB


def test_mp2_uhf_closed_shell_equals_rhf(oracle):
    mol, bas = g2("h2o_q0m1", "631g_star")
    _, Cm, e = _rhf(oracle, mol, bas, 1e-10)
    r = oracle.mp2_rhf(bas.ptr, bas.n, 5, 1, e, Cm, 0.0)
    u = oracle.mp2_uhf(bas.ptr, bas.n, 5, 5, 1, e, e, Cm, Cm, 0.0)
    assert abs(r - u) < 1e-14


def test_mp2_uhf_oracle_vs_reference(oracle, reference_mp2):
    mol, bas, CA, CB, eA, eB = _water_cation(oracle)
    rules = oracle.mp2_uhf(bas.ptr, bas.n, 5, 4, 1, eA, eB, CA, CB, 1e-8)
    ref = reference_mp2.mp2_uhf(bas.ptr, bas.n, mol.ptr, 5, 4, eA, eB, CA, CB)
    print(f"rules {rules:.15f} reference {ref:.15f}")
    assert abs(rules - ref) < 1e-14


def test_mp2_uhf_vs_textbook_spin_orbitals(oracle):
    """E2 = 1/4 sum |<ij||ab>|^2 / D per spin plus the alpha-beta sum, from
    the full AO tensor."""
    mol, bas, CA, CB, eA, eB = _water_cation(oracle)
    n = bas.n
    V = oracle.eri_pairs(bas.ptr, n)
    idx = np.array([[max(p, q) * (max(p, q) + 1) // 2 + min(p, q) for q in range(n)] for p in range(n)])
    T = V[idx[:, :, None, None], idx[None, None, :, :]]

    def mo(c1, c2, c3, c4):
        return np.einsum("pqrs,pa,qi,rb,sj->aibj", T, c1, c2, c3, c4, optimize=True)

    def denom(eo1, ev1, eo2, ev2):
        return eo1[:, None, None, None] - ev1[None, :, None, None] + eo2[None, None, :, None] - ev2[None, None, None, :]

    Ca, Cb = CA.T, CB.T
    oa, va, ob, vb = Ca[:, 1:5], Ca[:, 5:], Cb[:, 1:4], Cb[:, 4:]
    eo_a, ev_a, eo_b, ev_b = eA[1:5], eA[5:], eB[1:4], eB[4:]
    total = 0.0
    for o_, v_, eo, ev in ((oa, va, eo_a, ev_a), (ob, vb, eo_b, ev_b)):
        g = mo(o_, v_, o_, v_)
        total += 0.25 * np.sum((g - g.transpose(0, 3, 2, 1)) ** 2 / denom(eo, ev, eo, ev))
    total += np.sum(mo(oa, va, ob, vb) ** 2 / denom(eo_a, ev_a, eo_b, ev_b))
    x = oracle.mp2_uhf(bas.ptr, n, 5, 4, 1, eA, eB, CA, CB, 0.0)
    print(f"oracle {x:.15f} textbook {total:.15f}")
    assert abs(x - total) < 1e-13
