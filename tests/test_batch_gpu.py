"""Batched J/K builds (sqjk_batch_*): many molecules per class launch, one
HIP graph per plan.

- Every molecule's J, K_A, K_B from the batch equal its own context's
  build (per-molecule launches) to 1e-12: the batch changes only which
  launch evaluates a quartet, not the arithmetic.
- The graph is replayed several times with eager work and host syncs in
  between, and every replay is checked against the eager result.
- Two molecules of the batch are checked against the oracle (the CPU
  restatement of src/matrix.c:3773-3986) at 1e-10.
- Other tuning tables and mixed precision leave the results unchanged
  (within the mixed-precision bound for the latter).
Workload: sqamd.inputs.batch_instances (seeded small molecules, STO-3G,
6-31G and 6-31G*, a quarter UHF doublet cations).
"""
import numpy as np
import pytest

from sqamd import JKEngine, inputs
from sqamd.batch import JKBatch

pytestmark = pytest.mark.gpu
CUT = 1e-10


@pytest.fixture(scope="module")
def setup():
    inst = inputs.batch_instances(24, 77)
    bases = [inputs.make_basis(x["mol"], x["basis"]) for x in inst]
    engines = [JKEngine(b.ptr, b.n) for b in bases]
    dens = []
    for k, (x, b) in enumerate(zip(inst, bases)):
        rng = np.random.default_rng([77, k])
        A = rng.standard_normal((b.n, b.n)) * 0.05
        B = rng.standard_normal((b.n, b.n)) * 0.05
        dens.append((A + A.T, (B + B.T) if not x["restricted"] else A + A.T))
    ref = [e.build(PA, PB, CUT, CUT, x["restricted"]) for e, (PA, PB), x in zip(engines, dens, inst)]
    batch = JKBatch(engines, [x["restricted"] for x in inst])
    for k, (PA, PB) in enumerate(dens):
        batch.set_density(k, PA, PB)
    batch.set_cutoffs(CUT, CUT)
    yield inst, bases, engines, batch, dens, ref
    batch.close()
    for e in engines:
        e.close()


def _max_diff(batch, ref):
    d = 0.0
    for k, (J, KA, KB) in enumerate(ref):
        Jb, KAb, KBb = batch.result(k)
        d = max(d, np.abs(Jb - J).max(), np.abs(KAb - KA).max(), np.abs(KBb - KB).max())
    return d


def test_batch_eager_matches_per_molecule(setup):
    inst, bases, engines, batch, dens, ref = setup
    batch.build(graph=False)
    st = batch.stats()
    d = _max_diff(batch, ref)
    print(f"{len(inst)} molecules, n = {min(b.n for b in bases)}..{max(b.n for b in bases)}, "
          f"{sum(not x['restricted'] for x in inst)} UHF; {st['launches']} launches, "
          f"{st['quartets']} quartets, {st['kernel_ms']:.3f} ms; max diff {d:.2e}")
    assert d < 1e-12


def test_batch_graph_replays(setup):
    inst, bases, engines, batch, dens, ref = setup
    for rep in range(4):
        batch.build(graph=True)
        d = _max_diff(batch, ref)
        # eager work of this program between the replays
        e = engines[rep]
        PA, PB = dens[rep]
        e.build(PA, PB, CUT, CUT, inst[rep]["restricted"])
        print(f"replay {rep}: max diff {d:.2e} ({batch.stats()['kernel_ms']:.3f} ms)")
        assert d < 1e-12


def test_batch_against_oracle(setup, oracle):
    inst, bases, engines, batch, dens, ref = setup
    batch.build(graph=True)
    picks = [k for k in range(len(inst)) if bases[k].n <= 40][:2]
    assert picks
    for k in picks:
        b = bases[k]
        PA, PB = dens[k]
        sch = oracle.schwarz(b.ptr, b.n)
        Jo, KAo, KBo = oracle.jk(oracle.jk_context(False, 0), b.ptr, b.n, PA, PB, sch, CUT, CUT,
                                 inst[k]["restricted"])
        J, KA, KB = batch.result(k)
        d = max(np.abs(J - Jo).max(), np.abs(KA - KAo).max(), np.abs(KB - KBo).max())
        print(f"molecule {k} ({inst[k]['name']}/{inst[k]['basis']}, n={b.n}): vs oracle {d:.2e}")
        assert d < 1e-10


def test_batch_tuning_and_mixed(setup):
    inst, bases, engines, batch, dens, ref = setup
    rows = np.full((6, 6), 3)
    split = np.full((6, 6), 200)
    grid = np.full((6, 6), 1)
    batch.set_tuning(rows, split, grid)
    batch.build(graph=True)
    batch.build(graph=True)
    d = _max_diff(batch, ref)
    print(f"other tuning: max diff {d:.2e}")
    assert d < 1e-12
    batch.set_tuning()  # defaults
    batch.set_mixed(1e-4)
    batch.build(graph=True)
    batch.build(graph=True)
    dm = _max_diff(batch, ref)
    batch.set_mixed(0.0)
    print(f"mixed 1e-4: max diff {dm:.2e}")
    assert dm < 1e-7
