"""Checkpoint and density files (host C, no GPU): the reference's text
formats (src/check.c:432-570, src/basis.c:37-203, src/uhf.c:1414-1462,
1843-1870) written and read back, and the orbital projection of
src/check.c:306-430 checked against what the mathematics fixes: projecting
orbitals onto the basis they were stored in returns them unchanged, and a
projection onto a larger basis keeps a stored orbital's overlap with itself.
"""
import ctypes as C

import numpy as np

from molecules import g2
from sqamd import GTOBasis, Molecule, default_option, dptr, lib

PD = C.POINTER(C.c_double)


def _setup(L):
    L.save_checkpoint.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_double, C.c_double,
                                  C.c_double, PD, PD, PD, PD, C.c_void_p]
    L.load_checkpoint_orbital.argtypes = [C.c_int, PD, PD, PD, PD, C.c_void_p]
    L.load_checkpoint_orbital.restype = C.c_double
    L.load_checkpoint_molecule.argtypes = [C.c_void_p]
    L.load_checkpoint_molecule.restype = C.POINTER(Molecule)
    L.load_checkpoint_basis.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
    L.load_checkpoint_basis.restype = C.POINTER(GTOBasis)
    L.load_checkpoint_multiplicity.argtypes = [C.c_void_p]
    L.guess_checkpoint_orbital.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_void_p, PD,
                                           PD, PD]
    L.GTO_overlap.argtypes = [C.c_int, C.c_int, C.POINTER(GTOBasis)]
    L.GTO_overlap.restype = C.c_double
    L.sq_save_density.argtypes = [C.c_int, C.c_char_p, PD, PD]
    L.sq_load_density.argtypes = [C.c_int, C.c_int, C.c_char_p, PD, PD]


def _overlap(L, bas):
    n = bas.n
    S = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            S[i, j] = S[j, i] = L.GTO_overlap(i, j, bas.ptr)
    return S


def _orbitals(S, seed):
    """S-orthonormal orbitals, orbital-major (C[k, i])."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal(S.shape)
    w, V = np.linalg.eigh(S)
    X = V @ np.diag(w ** -0.5) @ V.T  # S^-1/2
    Q, _ = np.linalg.qr(A)
    return (X @ Q).T.copy()


def test_checkpoint_round_trip(tmp_path):
    L = lib()
    _setup(L)
    mol, bas = g2("h2o_q0m1", "631g_star")
    n = bas.n
    S = _overlap(L, bas)
    CA, CB = _orbitals(S, 1), _orbitals(S, 2)
    eA, eB = np.sort(np.random.default_rng(3).standard_normal(n)), np.sort(np.random.default_rng(4).standard_normal(n))
    opt = default_option(multiplicity=1)
    opt.CheckFile = str(tmp_path / "h2o.chk").encode()
    L.save_checkpoint(n, bas.ptr, mol.ptr, -76.0097959764, 1.5e-9, 2.5e-9, dptr(CA), dptr(CB), dptr(eA), dptr(eB),
                      C.byref(opt))
    text = open(tmp_path / "h2o.chk").read().split()
    assert text[0] == "SIAM_QUANTUM_CHECKPOINT" and "SECTION_ELECTRONIC_STRUCTURE_END" in text
    assert L.load_checkpoint_multiplicity(C.byref(opt)) == 1
    m = L.load_checkpoint_molecule(C.byref(opt)).contents
    assert m.nAtom == mol.s.nAtom and m.Q == 0
    assert all(m.Z[i] == mol.s.Z[i] and abs(m.x[i] - mol.s.x[i]) < 1e-7 for i in range(m.nAtom))
    nb = C.c_int()
    g = L.load_checkpoint_basis(C.byref(opt), C.byref(nb))
    assert nb.value == n
    for i in range(n):
        a, b = g[i], bas.arr[i]
        assert (a.l, a.m, a.n, a.nContract) == (b.l, b.m, b.n, b.nContract)
        for k in range(a.nContract):
            assert abs(a.exp[k] / b.exp[k] - 1) < 1e-8 and abs(a.coef[k] - b.coef[k]) < 1e-8
            assert abs(a.norm[k] / b.norm[k] - 1) < 1e-7  # recomputed 1/sqrt(<g|g>)
    cA, cB, fA, fB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
    E = L.load_checkpoint_orbital(n, dptr(cA), dptr(cB), dptr(fA), dptr(fB), C.byref(opt))
    assert abs(E + 76.0097959764) < 1e-8
    for x, y in ((cA, CA), (cB, CB), (fA, eA), (fB, eB)):
        assert np.allclose(x, y, rtol=1e-8, atol=1e-15)


def test_projection_identity_and_larger_basis(tmp_path):
    L = lib()
    _setup(L)
    mol, small = g2("h2o_q0m1", "sto3g")
    _, big = g2("h2o_q0m1", "631g_star")
    Ss = _overlap(L, small)
    CA = _orbitals(Ss, 5)
    z = np.zeros(small.n)
    opt = default_option(multiplicity=1)
    opt.CheckFile = str(tmp_path / "s.chk").encode()
    L.save_checkpoint(small.n, small.ptr, mol.ptr, -74.96, 0.0, 0.0, dptr(CA), dptr(CA), dptr(z), dptr(z), C.byref(opt))
    # same basis: the projection returns the stored orbitals
    PA, PB = np.zeros((small.n, small.n)), np.zeros((small.n, small.n))
    L.guess_checkpoint_orbital(small.n, small.ptr, mol.ptr, C.byref(opt), dptr(Ss), dptr(PA), dptr(PB))
    assert np.abs(PA - CA).max() < 1e-7 and np.abs(PB - CA).max() < 1e-7
    # larger basis: X = S^-1 Y C is the least-squares fit, so <X_k|X_k> <= 1
    # and <X_k|phi_k> = <X_k|X_k> (the residual is orthogonal to the span)
    Sb = _overlap(L, big)
    XA, XB = np.zeros((big.n, big.n)), np.zeros((big.n, big.n))
    L.guess_checkpoint_orbital(big.n, big.ptr, mol.ptr, C.byref(opt), dptr(Sb), dptr(XA), dptr(XB))
    norms = np.einsum("ki,ij,kj->k", XA[:small.n], Sb, XA[:small.n])
    assert np.all(norms <= 1 + 1e-8) and np.all(norms > 0.5)


def test_density_file_round_trip(tmp_path):
    L = lib()
    _setup(L)
    n = 9
    rng = np.random.default_rng(7)
    PA, PB = rng.standard_normal((n, n)), rng.standard_normal((n, n))
    path = str(tmp_path / "d.txt").encode()
    L.sq_save_density(n, path, dptr(PA), dptr(PB))
    QA, QB = np.zeros((n, n)), np.zeros((n, n))
    L.sq_load_density(n, 0, path, dptr(QA), dptr(QB))
    assert np.allclose(QA, PA, rtol=1e-8) and np.allclose(QB, PB, rtol=1e-8)
    L.sq_load_density(n, 1, path, dptr(QA), dptr(QB))  # restricted: beta = alpha
    assert np.array_equal(QA, QB)


def _ref_check():
    import pytest
    from oracle import ReferenceCheck
    try:
        return ReferenceCheck()
    except FileNotFoundError:
        pytest.skip("oracle/_ref/libsqref_check.so not built")


def _compare_loaded(Lr, opt, mol, bas, CA, CB, eA, eB, E):
    """Reads a checkpoint through loader library Lr and compares it with the
    data it was written from."""
    n = bas.n
    assert Lr.load_checkpoint_multiplicity(C.byref(opt)) == 2
    m = Lr.load_checkpoint_molecule(C.byref(opt)).contents
    assert m.nAtom == mol.s.nAtom and m.Q == mol.s.Q
    assert all(m.Z[i] == mol.s.Z[i] and abs(m.x[i] - mol.s.x[i]) < 1e-7 and abs(m.z[i] - mol.s.z[i]) < 1e-7
               for i in range(m.nAtom))
    nb = C.c_int()
    g = Lr.load_checkpoint_basis(C.byref(opt), C.byref(nb))
    assert nb.value == n
    for i in range(n):
        a, b = g[i], bas.arr[i]
        assert (a.l, a.m, a.n, a.nContract) == (b.l, b.m, b.n, b.nContract)
        assert abs(a.x0 - b.x0) < 1e-7 and abs(a.y0 - b.y0) < 1e-7 and abs(a.z0 - b.z0) < 1e-7
        for k in range(a.nContract):
            assert abs(a.exp[k] / b.exp[k] - 1) < 1e-8 and abs(a.coef[k] - b.coef[k]) < 1e-8
            assert abs(a.norm[k] / b.norm[k] - 1) < 1e-7
    cA, cB, fA, fB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
    E2 = Lr.load_checkpoint_orbital(n, dptr(cA), dptr(cB), dptr(fA), dptr(fB), C.byref(opt))
    assert abs(E2 - E) < 1e-8
    for x, y in ((cA, CA), (cB, CB), (fA, eA), (fB, eB)):
        assert np.allclose(x, y, rtol=1e-7, atol=1e-12)


def test_checkpoint_moves_between_programs(tmp_path):
    """A checkpoint written here loads in the reference's own reader
    (src/check.c:37-300, compiled into oracle/_ref), and one written by the
    reference's save_checkpoint (src/check.c:432-570) loads here: molecule,
    multiplicity, basis and orbitals of an L-shell (6-31G*) UHF doublet."""
    L = lib()
    _setup(L)
    R = _ref_check().L
    mol, bas = g2("nh3_q0m1", "631g_star", charge=1)
    n = bas.n
    S = _overlap(L, bas)
    CA, CB = _orbitals(S, 11), _orbitals(S, 12)
    eA = np.sort(np.random.default_rng(13).standard_normal(n))
    eB = np.sort(np.random.default_rng(14).standard_normal(n))
    E = -55.8734520312
    for k, (writer, reader) in enumerate(((L, R), (R, L))):
        opt = default_option(multiplicity=2, restricted=0)
        opt.CheckFile = str(tmp_path / f"nh3_{k}.chk").encode()
        writer.save_checkpoint(n, bas.ptr, mol.ptr, E, 2e-9, 3e-9, dptr(CA), dptr(CB), dptr(eA), dptr(eB),
                               C.byref(opt))
        _compare_loaded(reader, opt, mol, bas, CA, CB, eA, eB, E)
