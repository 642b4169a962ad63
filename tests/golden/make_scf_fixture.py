"""Writes tests/golden/scf_w16.json: the converged RHF/6-31G* energy of the
(H2O)16 cluster (seed 1601, 304 functions) on the reference's own two-electron
code, and the frozen-core MP2 energy of those orbitals three ways: by the
reference's own mp2_rhf_aqij_Parallel (src/mp2.c:713-927,
oracle/_ref/libsqref_mp2.so, one process, one cycle: maxMem 9000 MB, so no
multipole substitution), and by the oracle's restatement of it with the
reference's 1e-8 screening rules and with none (the exact energy).

J and K come from the compiled reference (oracle/_ref/libsqref.so,
GTO_JK_Matrix_Quartet_Parallel of src/matrix.c:3773-3986) at near-exact
settings: fixedCutoff 1e-14 (the density-weighted screen and the multipole
substitution act only below it; fixedCutoff 0 would not do, since the
reference's memory cut-off search doubles fixedCutoff until the integral
bank fits, src/matrix.c:3703-3761, and never leaves 0), SCFCutoff 1e-13 (the
Schwarz pre-screen), a 100 MB integral bank per process, run as nCPU = 8
childID slices, each in its own process for the whole run (the reference
keeps per-process state between calls: its shells, multipole data and
integral bank belong to the childID of the first call).  S, T, V come from the
reference's GTO_overlap / GTO_kinetic / GTO_nuclei.  The SCF loop around them
is a plain restricted Hartree-Fock iteration (core-Hamiltonian guess, Pulay
DIIS on the Fock matrix, generalized eigenproblem by LAPACK through scipy)
converged to |dE| < 1e-11 Eh and a DIIS error below 1e-8: at convergence the
energy does not depend on how the iteration got there, so it pins the
device SCF's energy (its own cut-offs stay far below the 1e-8 Eh bar of the
test that reads this file).  The MP2 energy is not stationary in the
orbitals; at this convergence it is fixed to a few 1e-10 Eh.  The
reference's own 1e-8 screening rules (MP2_SCHWARZ_CUTOFF) move it by
1.2e-8 Eh here, which is why the exact energy is written beside it.

Usage (this container: the SCF about 2 minutes on 8 cores, the three MP2
energies about 35 minutes, the exact one on the oracle the longest):
  python tests/golden/make_scf_fixture.py
"""
import ctypes as C
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "siam-quantum_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
NPROC = 8

_w = {}


def _init():
    from molecules import cluster
    from oracle import Reference
    from sqamd import default_option
    mol, bas = cluster(16, 1601)
    ref = Reference()
    _w.update(mol=mol, bas=bas, ref=ref, sch=ref.schwarz(bas.ptr, bas.n),
              opt=default_option(SCFCutoff=1e-13, restricted=1, maxMem=100, nCPU=NPROC))


def _jk_slice(args):
    child, P = args
    from sqamd.abi import dptr
    bas, ref, opt = _w["bas"], _w["ref"], _w["opt"]
    n = bas.n
    J, KA, KB = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    ref.L.GTO_JK_Matrix_Quartet_Parallel(child, n, dptr(P), dptr(P), bas.ptr, dptr(_w["sch"]), 1e-14,
                                         dptr(J), dptr(KA), dptr(KB), C.byref(opt))
    return J, KA


def _mp2_job(args):
    kind, nocc, eps, Cr = args
    from molecules import cluster
    from oracle import Oracle, ReferenceMP2
    mol, bas = cluster(16, 1601)
    if kind == "reference":
        return ReferenceMP2().mp2(bas.ptr, bas.n, mol.ptr, nocc, eps, Cr, max_mem=9000)
    o = Oracle()
    return o.mp2_rhf(bas.ptr, bas.n, nocc, o.mp2_frozen_core(mol.ptr), eps, Cr, 1e-8 if kind == "rules" else 0.0)


def main():
    import scipy.linalg as sl
    from molecules import cluster
    from oracle import Reference
    mol, bas = cluster(16, 1601)
    ref = Reference()
    n, ne = bas.n, mol.n_electrons
    nocc = ne // 2
    t0 = time.time()
    S, H = np.zeros((n, n)), np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            S[i, j] = S[j, i] = ref.L.GTO_overlap(i, j, bas.ptr)
            H[i, j] = H[j, i] = ref.L.GTO_kinetic(i, j, bas.ptr) + ref.L.GTO_nuclei(i, j, bas.ptr, mol.ptr)
    m = mol.ptr.contents
    xyz = np.array([[m.x[i], m.y[i], m.z[i]] for i in range(m.nAtom)])
    Z = np.array([m.Z[i] for i in range(m.nAtom)], dtype=float)
    Enn = sum(Z[i] * Z[j] / np.linalg.norm(xyz[i] - xyz[j]) for i in range(len(Z)) for j in range(i))
    eps, Cm = sl.eigh(H, S)
    P = Cm[:, :nocc] @ Cm[:, :nocc].T  # alpha density
    E, Fs, Es, it = None, [], [], 0
    pools = [mp.Pool(1, initializer=_init) for _ in range(NPROC)]  # one process per childID
    try:
        while True:
            it += 1
            jobs = [pools[c].apply_async(_jk_slice, ((c, P),)) for c in range(NPROC)]
            parts = [j.get() for j in jobs]
            JT = sum(p[0] for p in parts)
            KA = sum(p[1] for p in parts)
            F = H + JT + KA  # KA holds -K[P_alpha]
            Eold, E = E, Enn + float(np.sum(P * (H + F)))
            err = F @ P @ S - S @ P @ F
            Fs.append(F)
            Es.append(err)
            Fs, Es = Fs[-8:], Es[-8:]
            k = len(Fs)
            B = -np.ones((k + 1, k + 1))
            B[k, k] = 0.0
            for a in range(k):
                for b in range(k):
                    B[a, b] = np.sum(Es[a] * Es[b])
            rhs = np.zeros(k + 1)
            rhs[k] = -1.0
            c = np.linalg.solve(B, rhs)[:k]
            Fd = sum(ci * Fi for ci, Fi in zip(c, Fs))
            eps, Cm = sl.eigh(Fd, S)
            P = Cm[:, :nocc] @ Cm[:, :nocc].T
            emax = float(np.abs(err).max())
            print(f"iter {it}: E = {E:.12f}  max|FPS-SPF| = {emax:.2e}  {time.time() - t0:.0f} s", flush=True)
            if Eold is not None and abs(E - Eold) < 1e-11 and emax < 1e-8:
                break
            if it >= 60:
                raise RuntimeError("SCF did not converge")
    finally:
        for p in pools:
            p.terminate()
    t_scf = time.time() - t0
    from oracle import ReferenceMP2
    ncore = ReferenceMP2().L.getFrozenCore_RHF(mol.ptr)
    Cr = np.ascontiguousarray(Cm.T)
    with mp.Pool(3) as pool:
        e2, e2_rules, e2_exact = pool.map(_mp2_job, [(k, nocc, eps, Cr) for k in ("reference", "rules", "exact")])
    out = {"workload": "(H2O)16 cluster, seed 1601, RHF/6-31G*", "n": n, "electrons": ne,
           "energy": E, "iterations": it, "converged": "|dE| < 1e-11, max|FPS-SPF| < 1e-8",
           "two_electron": "reference GTO_JK_Matrix_Quartet_Parallel, fixedCutoff 1e-14, SCFCutoff 1e-13",
           "seconds": round(t_scf, 1),
           "mp2_energy": e2, "mp2_frozen_core": ncore,
           "mp2": "reference mp2_rhf_aqij_Parallel on the final orbitals, nCPU 1, one cycle",
           "mp2_energy_oracle_rules": e2_rules, "mp2_energy_oracle_exact": e2_exact,
           "mp2_seconds": round(time.time() - t0 - t_scf, 1)}
    with open(os.path.join(HERE, "scf_w16.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out)


if __name__ == "__main__":
    main()
