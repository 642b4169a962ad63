"""Shared test inputs: molecules and bases built through the product's input
path (sqamd.inputs) so that oracle and engine see the same GTOBasis_t array."""
import os

import numpy as np

from sqamd import inputs

G2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "g2")


def water(basis="sto3g"):
    Z, X = inputs.water_c1()
    mol = inputs.mol_from_angstrom(Z, X)
    return mol, inputs.make_basis(mol, basis)


def g2(name, basis, charge=0):
    mol = inputs.read_xyz(os.path.join(G2, name + ".xyz"), charge)
    return mol, inputs.make_basis(mol, basis)


def cluster(nmol, seed, basis="631g_star"):
    Z, X = inputs.water_cluster(nmol, seed)
    mol = inputs.mol_from_angstrom(Z, X)
    return mol, inputs.make_basis(mol, basis)


def random_density(n, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) * scale
    return A + A.T
