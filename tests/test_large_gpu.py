"""J/K of the benchmark-size systems against the reference's own J/K.

Fixtures tests/golden/jk_<name>.npz hold, for a seeded density, 8 full rows,
the diagonals, J @ v and K @ v for a seeded v, and Tr(PJ), Tr(PK) of the
reference's GTO_JK_Matrix_Quartet_Parallel (src/matrix.c:3773-3986, built
from the reference sources, 8 processes on the PQ % 8 slices; see
tests/golden/make_jk_fixtures.py).  Workloads:
  w16  (H2O)16 6-31G*, 304 functions, cut-off 1e-12
  w32  (H2O)32 6-31G*, 608 functions, cut-off 1e-12
  c60  C60 6-31G*, 900 functions, cut-off 1e-10
Bar (plain FP64 path): 1e-9 max-abs on every stored element; J @ v, K @ v
sum n elements and get 1e-8.
"""
import os

import numpy as np
import pytest

from molecules import random_density
from sqamd import JKEngine, inputs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _molecule(name):
    if name == "c60":
        Z, X = inputs.c60()
    else:
        nm = int(name[1:])
        Z, X = inputs.water_cluster(nm, {16: 1601, 32: 3201}[nm])
    mol = inputs.mol_from_angstrom(Z, X)
    return mol, inputs.make_basis(mol, "631g_star")


@pytest.mark.parametrize("name", ["w16", "w32", "c60"])
def test_jk_matches_reference_fixture(name):
    fx = np.load(os.path.join(GOLD, f"jk_{name}.npz"))
    mol, bas = _molecule(name)
    n = bas.n
    assert n == int(fx["n"])
    cut = float(fx["cutoff"])
    P = random_density(n, int(fx["seed"]), scale=float(fx["scale"]))
    eng = JKEngine(bas.ptr, n)
    J, K, _ = eng.build(P, P, cut, cut, True)
    st = eng.stats()
    work = eng.class_work()
    eng.close()
    rows = fx["rows"]
    d = {
        "J rows": np.abs(J[rows] - fx["J_rows"]).max(),
        "K rows": np.abs(K[rows] - fx["K_rows"]).max(),
        "J diag": np.abs(np.diag(J) - fx["Jdiag"]).max(),
        "K diag": np.abs(np.diag(K) - fx["Kdiag"]).max(),
        "Jv": np.abs(J @ fx["v"] - fx["Jv"]).max(),
        "Kv": np.abs(K @ fx["v"] - fx["Kv"]).max(),
    }
    trJ, trK = np.sum(P * J), np.sum(P * K)
    print(f"{name}: n={n} quartets={st['quartets']} surviving={sum(w[1] for w in work)} "
          f"kernel={st['kernel_ms']:.2f} ms  " + "  ".join(f"{k}={v:.2e}" for k, v in d.items()) +
          f"  Tr(PJ) {trJ:.12e} vs {float(fx['trPJ']):.12e}  Tr(PK) {trK:.12e} vs {float(fx['trPK']):.12e}",
          flush=True)
    for k in ("J rows", "K rows", "J diag", "K diag"):
        assert d[k] < 1e-9, k
    assert d["Jv"] < 1e-8 and d["Kv"] < 1e-8
    assert abs(trJ - float(fx["trPJ"])) < 1e-9 * max(1.0, abs(trJ))
    assert abs(trK - float(fx["trPK"])) < 1e-9 * max(1.0, abs(trK))


def test_class_work_counts():
    """Surviving work: with no density-weighted cut every Schwarz-screened
    quartet survives, so the counted quartets equal the plan's; the primitive
    cut only removes primitive quartets; a density cut removes quartets."""
    mol, bas = _molecule("w16")
    n = bas.n
    P = random_density(n, 3, scale=0.01)
    eng = JKEngine(bas.ptr, n)
    eng.build(P, P, 0.0, 1e-12, True)
    plan = eng.class_stats()
    work = eng.class_work()
    for (lq, q, p), (lq2, wq, wp, wq32, wp32) in zip(plan, work):
        assert lq == lq2
        assert wq == q and wq32 == 0 and wp32 == 0, lq
        assert 0 < wp <= p or q == 0, lq
    eng.build(P, P, 1e-8, 1e-12, True)
    work2 = eng.class_work()
    eng.close()
    assert sum(w[1] for w in work2) < sum(w[1] for w in work)
