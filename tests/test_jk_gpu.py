"""J/K parity of the HIP engine against the oracle (CPU restatement of
src/matrix.c:3773-3986 and src/int.c), on seeded densities.

Tolerances: the plain FP64 path must reproduce the reference Fock pieces to
1e-9 max-abs; the cases here hold 1e-10.
"""
import os

import numpy as np
import pytest

from molecules import G2, cluster, g2, random_density, water
from sqamd import JKEngine, default_option, dptr, inputs, lib

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _compare(oracle, mol, bas, restricted, fixed_cut=0.0, scf_cut=0.0, seed=1):
    n = bas.n
    sch = oracle.schwarz(bas.ptr, n)
    PA = random_density(n, seed)
    PB = PA if restricted else random_density(n, seed + 100)
    ctx = oracle.jk_context(False, 0)
    Jo, KAo, KBo = oracle.jk(ctx, bas.ptr, n, PA, PB, sch, fixed_cut, scf_cut, restricted)
    eng = JKEngine(bas.ptr, n, schwarz=sch)
    Jg, KAg, KBg = eng.build(PA, PB, fixed_cut, scf_cut, restricted)
    st = eng.stats()
    eng.close()
    print(f"n={n} restricted={restricted} cut={fixed_cut:g} quartets={st['quartets']} "
          f"kernel={st['kernel_ms']:.3f}ms dJ={np.abs(Jg - Jo).max():.2e} "
          f"dKA={np.abs(KAg - KAo).max():.2e} dKB={np.abs(KBg - KBo).max():.2e}", flush=True)
    assert np.abs(Jg - Jo).max() < TOL
    assert np.abs(KAg - KAo).max() < TOL
    assert np.abs(KBg - KBo).max() < TOL


@pytest.mark.parametrize("restricted", [True, False])
def test_water_sto3g(oracle, restricted):
    mol, bas = water("sto3g")
    _compare(oracle, mol, bas, restricted)


@pytest.mark.parametrize("restricted", [True, False])
def test_water_631gs(oracle, restricted):
    mol, bas = g2("h2o_q0m1", "631g_star")
    _compare(oracle, mol, bas, restricted)


def test_ethylene_631gs_uhf(oracle):
    mol, bas = g2("c2h4_q0m1", "631g_star")
    _compare(oracle, mol, bas, False, seed=7)


def test_screened_build_matches(oracle):
    """Reference screening (Schwarz x density, 3-step cutoff) on a dimer."""
    mol, bas = cluster(2, 11, "631g_star")
    _compare(oracle, mol, bas, True, fixed_cut=1e-6, scf_cut=1e-10, seed=3)


def test_schwarz_matches(oracle):
    mol, bas = g2("ch2o_q0m1", "631g_star")
    so = oracle.schwarz(bas.ptr, bas.n)
    eng = JKEngine(bas.ptr, bas.n)
    sg = eng.schwarz()
    eng.close()
    assert np.abs(sg - so).max() < 1e-11


def test_dropin_accumulates_and_resets(oracle):
    """GTO_JK_Matrix_Quartet_Parallel keeps the reference semantics: it adds
    into JT/KA/KB, copies KA to KB when restricted, and a call with nBasis=0
    drops the state (src/matrix.c:3817-3828, 3974-3982)."""
    L = lib()
    mol, bas = water("sto3g")
    n = bas.n
    sch = oracle.schwarz(bas.ptr, n)
    P = random_density(n, 5)
    opt = default_option(SCFCutoff=0.0, restricted=1)
    J, KA, KB = np.ones((n, n)), np.ones((n, n)), np.zeros((n, n))
    L.GTO_JK_Matrix_Quartet_Parallel(0, n, dptr(P), dptr(P), bas.ptr, dptr(sch), 0.0,
                                     dptr(J), dptr(KA), dptr(KB), opt)
    z = np.zeros(1)
    L.GTO_JK_Matrix_Quartet_Parallel(0, 0, dptr(z), dptr(z), bas.ptr, dptr(z), 0.0,
                                     dptr(z), dptr(z), dptr(z), opt)
    ctx = oracle.jk_context(False, 0)
    Jo, KAo, _ = oracle.jk(ctx, bas.ptr, n, P, P, sch, 0.0, 0.0, True)
    assert np.abs(J - (Jo + 1.0)).max() < TOL
    assert np.abs(KA - (KAo + 1.0)).max() < TOL
    assert np.array_equal(KA, KB)


D3 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "basis", "d3_test.txt")


@pytest.mark.parametrize("restricted", [True, False])
def test_contracted_d_split_tiles(oracle, restricted):
    """Three-primitive d shells (a synthetic GAMESS basis, tests/golden/basis):
    a (dd|dd) quartet then has 81 primitive quartets, above the split limit of
    its class, so its tiles are cut along the ket primitive pairs; both the
    cooperative (dd|dd) kernel and the per-thread d classes must sum each
    primitive range once."""
    mol = inputs.read_xyz(os.path.join(G2, "ch2o_q0m1.xyz"))
    bas = inputs.make_basis(mol, D3)
    _compare(oracle, mol, bas, restricted, seed=4)
