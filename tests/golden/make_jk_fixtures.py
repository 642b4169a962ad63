"""Writes the large-system J/K parity fixtures tests/golden/jk_<name>.npz.

For each workload the reference's own J/K driver, GTO_JK_Matrix_Quartet_Parallel
(src/matrix.c:3773-3986, compiled from the reference sources into
oracle/_ref/libsqref.so by oracle/Makefile), builds J and K on a seeded
symmetric density with the screening cut-off used by the benchmark.  The
build is split the way the reference splits it over processes: child k of
nCPU = 8 takes the bra shell pairs PQ % 8 == k (src/matrix.c:3909-3910) in a
process of its own, and the eight partial results are summed.  The first
call after a reset uses exact integrals only (no monopole substitution
before the error bank exists, src/matrix.c:3946-3955).

Stored per workload (compact; the full matrices are 6.5 MB at n = 900):
  rows, J_rows, K_rows   8 seeded full rows of J and K,
  v, Jv, Kv              J @ v and K @ v for a seeded vector v (every element
                         of J and K enters),
  Jdiag, Kdiag           the diagonals,
  trPJ, trPK             Tr(P J), Tr(P K),
plus n, cutoff, density seed and scale.  The density is
tests/molecules.random_density(n, seed, scale), rebuilt by the test.

Usage (this container, with oracle/_ref built):
  python tests/golden/make_jk_fixtures.py [w16 w32 c60]
"""
import ctypes as C
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

NCPU = 8
# name: (geometry, density seed, density scale, cut-off)
WORKLOADS = {
    "w16": ("water16", 1616, 0.01, 1e-12),
    "w32": ("water32", 3232, 0.01, 1e-12),
    "c60": ("c60", 6060, 0.01, 1e-10),
}


def molecule(geom):
    from sqamd import inputs
    if geom == "c60":
        Z, X = inputs.c60()
    else:
        nm = int(geom[5:])
        Z, X = inputs.water_cluster(nm, {16: 1601, 32: 3201}[nm])
    mol = inputs.mol_from_angstrom(Z, X)
    return mol, inputs.make_basis(mol, "631g_star")


def child(args):
    name, k = args
    from oracle import Reference, default_option
    from sqamd import dptr
    from molecules import random_density
    geom, seed, scale, cut = WORKLOADS[name]
    mol, bas = molecule(geom)
    n = bas.n
    ref = Reference()
    sch = ref.schwarz(bas.ptr, n)
    P = random_density(n, seed, scale=scale)
    opt = default_option(SCFCutoff=cut, restricted=1, maxMem=1, nCPU=NCPU)
    ref.jk_reset(opt)
    J, KA, KB = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    t0 = time.time()
    ref.L.GTO_JK_Matrix_Quartet_Parallel(k, n, dptr(P), dptr(P), bas.ptr, dptr(sch), cut,
                                         dptr(J), dptr(KA), dptr(KB), C.byref(opt))
    ref.jk_reset(opt)
    print(f"{name}: child {k} of {NCPU} done in {time.time() - t0:.1f} s", flush=True)
    return J, KA


def make(name):
    from molecules import random_density
    geom, seed, scale, cut = WORKLOADS[name]
    mol, bas = molecule(geom)
    n = bas.n
    with mp.get_context("spawn").Pool(NCPU) as pool:
        parts = pool.map(child, [(name, k) for k in range(NCPU)])
    J = sum(p[0] for p in parts)
    K = sum(p[1] for p in parts)
    P = random_density(n, seed, scale=scale)
    rng = np.random.default_rng(seed + 1)
    rows = np.sort(rng.choice(n, 8, replace=False))
    v = rng.standard_normal(n)
    out = os.path.join(HERE, f"jk_{name}.npz")
    np.savez_compressed(out, n=n, cutoff=cut, seed=seed, scale=scale, rows=rows,
                        J_rows=J[rows], K_rows=K[rows], v=v, Jv=J @ v, Kv=K @ v,
                        Jdiag=np.diag(J).copy(), Kdiag=np.diag(K).copy(),
                        trPJ=np.sum(P * J), trPK=np.sum(P * K))
    print(f"{name}: n={n} Tr(PJ)={np.sum(P * J):.12e} Tr(PK)={np.sum(P * K):.12e} -> {out}", flush=True)


if __name__ == "__main__":
    for name in (sys.argv[1:] or list(WORKLOADS)):
        make(name)
