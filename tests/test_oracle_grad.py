"""Pins of the oracle's force routines (oracle/sq_oracle.c, src/grad.c
restated) — CPU only.

Each piece is checked against the reference's own grad.c compiled into
oracle/_ref, and the assembled forces against central finite differences
of the oracle's converged SCF energy, which fix them independently of
either program (a wrong sign, a dropped term or a transposed index moves
the forces by far more than the 1e-6 Eh/bohr bar).
"""
import ctypes as C

import numpy as np
import pytest

from molecules import g2, random_density, water
from sqamd import inputs

BOHR = 1.8897261245650618


def basis_to_atom(mol, bas):
    m, g = mol.ptr.contents, bas.ptr
    out = []
    for p in range(bas.n):
        out.append(next(i for i in range(m.nAtom) if (g[p].x0, g[p].y0, g[p].z0) == (m.x[i], m.y[i], m.z[i])))
    return np.array(out, dtype=np.int32)


@pytest.mark.parametrize("kfactor", [1.0, 0.5, 0.0])
def test_grad_eri_vs_compiled_reference(oracle, reference_grad, kfactor):
    mol, bas = g2("h2o_q0m1", "631g_star")
    n, na = bas.n, mol.ptr.contents.nAtom
    b2a = basis_to_atom(mol, bas)
    PA, PB = random_density(n, 1, 0.2), random_density(n, 2, 0.2)
    G = oracle.grad_eri(bas.ptr, n, b2a, na, PA + PB, PA - PB, kfactor)
    R = reference_grad.grad_eri(bas.ptr, n, b2a, na, PA + PB, PA - PB, kfactor)
    assert np.abs(G - R).max() < 1e-12 * max(1.0, np.abs(R).max())
    # translational invariance of the two-electron energy
    assert np.abs(G.sum(0)).max() < 1e-12


def test_grad_eri_slices_sum(oracle, reference_grad):
    """childID/nCPU slices of the shell-pair list add up to the whole."""
    mol, bas = g2("c2h4_q0m1", "631g_star")
    n, na = bas.n, mol.ptr.contents.nAtom
    b2a = basis_to_atom(mol, bas)
    PA = random_density(n, 3, 0.1)
    whole = oracle.grad_eri(bas.ptr, n, b2a, na, 2 * PA, 0 * PA)
    parts = sum(oracle.grad_eri(bas.ptr, n, b2a, na, 2 * PA, 0 * PA, child=c, ncpu=3) for c in range(3))
    ref = sum(reference_grad.grad_eri(bas.ptr, n, b2a, na, 2 * PA, 0 * PA, child=c, ncpu=3) for c in range(3))
    assert np.abs(whole - parts).max() < 1e-12
    assert np.abs(parts - ref).max() < 1e-12


def test_one_electron_derivatives_vs_compiled_reference(oracle, reference_grad):
    mol, bas = g2("h2o_q0m1", "631g_star")
    g, n = bas.ptr, bas.n
    for d in range(3):
        for p in range(n):
            for q in range(n):
                for fo, fr in ((oracle.L.or_grad_overlap, reference_grad.L.GTO_grad_overlap),
                               (oracle.L.or_grad_kinetic, reference_grad.L.GTO_grad_kinetic)):
                    a, b = fo(p, q, g, d), fr(p, q, g, d)
                    assert abs(a - b) < 1e-12 * max(1.0, abs(b))
                a = oracle.L.or_grad_nai(p, q, g, mol.ptr, d)
                b = reference_grad.L.GTO_grad_nai(p, q, g, mol.ptr, d)
                assert abs(a - b) < 1e-12 * max(1.0, abs(b))
                a = oracle.L.or_grad_moment(p, q, g, 0, 1, 0, 0.1, 0.2, -0.3, d)
                b = reference_grad.L.GTO_grad_moment(p, q, g, 0, 1, 0, 0.1, 0.2, -0.3, d)
                assert abs(a - b) < 1e-12 * max(1.0, abs(b))


_EFI_SCRIPT = """
import json, sys, ctypes as C
sys.path[:0] = sys.argv[1:3]
from molecules import g2
from oracle import ReferenceGrad
r = ReferenceGrad()
mol, bas = g2("h2o_q0m1", "631g_star")
b = [C.c_double() for _ in range(3)]
out = []
for p in range(bas.n):
    for q in range(p + 1):
        r.L.GTO_efi(p, q, bas.ptr, 0.3, -0.2, 0.5, *[C.byref(v) for v in b])
        out.append([v.value for v in b])
print(json.dumps(out))
"""


def test_field_integral_vs_compiled_reference(oracle, reference_grad):
    """The reference's efi_1d reads CPx_pow[I-u-1] at index -1 with a zero
    factor (src/int.c:451-456): whatever the stack holds there is multiplied
    by zero, which gives NaN when it holds an infinity or a NaN — in a
    process that has run other code first, every value can come back NaN.
    The compiled reference therefore runs in a fresh process; values it
    still returns as non-finite are skipped, the rest must agree and most
    must be finite."""
    import json
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    env = dict(os.environ, PYTHONPATH=os.path.join(root, "siam-quantum_amd"))
    res = subprocess.run([sys.executable, "-c", _EFI_SCRIPT, here, os.path.join(root, "oracle")],
                         capture_output=True, text=True, env=env, timeout=120)
    assert res.returncode == 0, res.stderr[-2000:]
    ref = json.loads(res.stdout.strip().splitlines()[-1])
    mol, bas = g2("h2o_q0m1", "631g_star")
    g, n = bas.ptr, bas.n
    a = [C.c_double() for _ in range(3)]
    compared = skipped = k = 0
    for p in range(n):
        for q in range(p + 1):
            oracle.L.or_gto_efi(p, q, g, 0.3, -0.2, 0.5, *[C.byref(v) for v in a])
            for u, v in zip(a, ref[k]):
                assert np.isfinite(u.value)
                if v is None or not np.isfinite(v):
                    skipped += 1
                    continue
                compared += 1
                assert abs(u.value - v) < 1e-12 * max(1.0, abs(v))
            k += 1
    assert compared > 10 * skipped


def test_field_integral_is_a_nuclear_derivative(oracle):
    """e = d/dC nai(C): central difference of the attraction integral."""
    mol, bas = water("sto3g")
    g = bas.ptr
    e = [C.c_double() for _ in range(3)]
    h = 1e-5
    for p, q in ((0, 0), (1, 0), (4, 2), (6, 5)):
        oracle.L.or_gto_efi(p, q, g, 0.3, -0.2, 0.5, *[C.byref(v) for v in e])
        for d in range(3):
            c0 = [0.3, -0.2, 0.5]
            cp, cm = list(c0), list(c0)
            cp[d] += h
            cm[d] -= h
            vals = []
            for c in (cp, cm):
                v = 0.0
                for a in range(g[p].nContract):
                    for b in range(g[q].nContract):
                        v += g[p].coef[a] * g[q].coef[b] * g[p].norm[a] * g[q].norm[b] * oracle.L.or_nai(
                            g[p].x0, g[p].y0, g[p].z0, 1.0, g[p].l, g[p].m, g[p].n, g[p].exp[a],
                            g[q].x0, g[q].y0, g[q].z0, 1.0, g[q].l, g[q].m, g[q].n, g[q].exp[b], *c)
                vals.append(v)
            assert abs((vals[0] - vals[1]) / (2 * h) - e[d].value) < 1e-8


def test_vnn_vs_compiled_reference(oracle, reference_grad):
    mol, _ = g2("c2h4_q0m1", "sto3g")
    na = mol.ptr.contents.nAtom
    a, b = np.zeros((3, na)), np.zeros((3, na))
    from sqamd.abi import dptr
    oracle.L.or_grad_vnn(mol.ptr, dptr(a[0]), dptr(a[1]), dptr(a[2]))
    reference_grad.L.GradVnn(mol.ptr, dptr(b[0]), dptr(b[1]), dptr(b[2]))
    assert np.abs(a - b).max() < 1e-14 * np.abs(b).max()


def _solve(oracle, Z, X, basis, charge=0, mult=1, orbitals=False):
    mol = inputs.mol_from_angstrom(Z, X, charge)
    bas = inputs.make_basis(mol, basis)
    ne = int(sum(Z)) - charge
    nEA = (ne + mult - 1) // 2
    res = oracle.uhf(bas.ptr, bas.n, mol.ptr, nEA, ne - nEA, restricted=(mult == 1), scf_conv=1e-11,
                     scf_max=200, three_step=False, use_mpole=False, max_mem=0, want_orbitals=orbitals)
    return res, mol, bas, nEA, ne - nEA


@pytest.mark.parametrize("basis,charge,mult", [("sto3g", 0, 1), ("sto3g", 1, 2), ("631g_star", 0, 1)])
def test_forces_vs_finite_differences(oracle, basis, charge, mult):
    """Water RHF and its UHF doublet cation: analytic forces = -dE/dR by
    central differences of the oracle SCF energy (h = 1e-4 bohr, SCF
    converged to 1e-11; truncation and SCF noise stay below 1e-7)."""
    Z, X = inputs.water_c1()
    (E, it, orbs), mol, bas, nEA, nEB = _solve(oracle, Z, X, basis, charge, mult, True)
    F = oracle.force(bas.ptr, bas.n, mol.ptr, nEA, nEB, orbs)
    assert np.abs(F.sum(0)).max() < 1e-10
    h = 1e-4 / BOHR
    coords = [(a, d) for a in range(3) for d in range(3)] if basis == "sto3g" else [(0, 1), (1, 0), (2, 1)]
    for a, d in coords:
        Xp, Xm = X.copy(), X.copy()
        Xp[a, d] += h
        Xm[a, d] -= h
        Ep = _solve(oracle, Z, Xp, basis, charge, mult)[0][0]
        Em = _solve(oracle, Z, Xm, basis, charge, mult)[0][0]
        assert abs(-(Ep - Em) / 2e-4 - F[a, d]) < 2e-7, (a, d)


def test_forces_in_a_field_vs_finite_differences(oracle):
    """The field terms: the SCF energy with a uniform field is not what the
    oracle SCF solves, so the pin here is the field part alone — a field
    changes the forces by -Z E on the nuclei plus 2 sum PT d<p|r|q>/dR E,
    which must be linear in E and vanish in sum over atoms for a neutral
    molecule only in the electronic + nuclear total."""
    Z, X = inputs.water_c1()
    (E, it, orbs), mol, bas, nEA, nEB = _solve(oracle, Z, X, "sto3g", 0, 1, True)
    F0 = oracle.force(bas.ptr, bas.n, mol.ptr, nEA, nEB, orbs)
    F1 = oracle.force(bas.ptr, bas.n, mol.ptr, nEA, nEB, orbs, field=(0.01, 0.0, 0.0))
    F2 = oracle.force(bas.ptr, bas.n, mol.ptr, nEA, nEB, orbs, field=(0.02, 0.0, 0.0))
    assert np.abs((F2 - F0) - 2 * (F1 - F0)).max() < 1e-12
    assert np.abs(F1 - F0).max() > 1e-4
