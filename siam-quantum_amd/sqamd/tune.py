"""Autotuner of the J/K work plan (sqjk_set_tuning / sqjk_batch_set_tuning).

Three knobs per class pair (bra pair class X, ket pair class Y; 6 x 6
tables): bra rows per tile, the integral-primitive budget above which a
segment's ket primitive pairs are split over tiles, and the grid's multiple
of the resident workgroups.  None changes the result, only the time.  The
tuner searches one knob at a time (rows, then split, then grid), trying each
candidate for every class pair at once and keeping, per class pair, the
candidate with the shortest kernel time (per-class HIP events of eager
builds, best of `reps`).  Class launches are independent, so one sweep
times all class pairs together.

  python -m sqamd.tune batch --out siam-quantum_amd/data/tune/batch_c4.json
  python -m sqamd.tune c60   --out siam-quantum_amd/data/tune/c60.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

CANDIDATES = {"rows": [1, 2, 3, 4], "split": [288, 576, 1152], "grid": [0, 2, 4, 8]}
KNOBS = ["rows", "split", "grid"]
DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data", "tune")


def class_pairs():
    """(X, Y) of the 21 classes in sqjk_class_stats order."""
    return [(X, Y) for X in range(6) for Y in range(X + 1)]


def tune(run, start, reps=2, log=print):
    """run(tables) -> per-class kernel ms (21, class order); start: dict of
    6x6 tables.  Returns (tables, report)."""
    cur = {k: np.array(start[k], dtype=np.int32).reshape(6, 6).copy() for k in KNOBS}
    t0 = np.array(run(cur))
    report = {"before_ms": float(t0.sum()), "stages": []}
    for knob in KNOBS:
        times = {}
        for v in CANDIDATES[knob]:
            trial = {k: cur[k].copy() for k in KNOBS}
            trial[knob][:, :] = v
            best = None
            for _ in range(reps):
                t = np.array(run(trial))
                best = t if best is None else np.minimum(best, t)
            times[v] = best
            log(f"tune {knob}={v}: {best.sum():.3f} ms")
        for k, (X, Y) in enumerate(class_pairs()):
            vals = {v: times[v][k] for v in times}
            if max(vals.values()) <= 0.0:
                continue
            cur[knob][X, Y] = min(vals, key=vals.get)
        report["stages"].append({"knob": knob, "sum_ms": {str(v): float(times[v].sum()) for v in times}})
    t1 = np.array(run(cur))
    report["after_ms"] = float(t1.sum())
    report["class_ms_after"] = [float(x) for x in t1]
    return cur, report


def save(path, tables, report, workload):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"workload": workload, "rows": tables["rows"].tolist(), "split": tables["split"].tolist(),
                   "grid": tables["grid"].tolist(), "report": report}, f, indent=1)


def load(path):
    d = json.load(open(path))
    return {k: np.array(d[k], dtype=np.int32) for k in KNOBS}


def _engine_tables(eng):
    import ctypes as C
    from .batch import _bind
    L = _bind()
    t = {k: np.zeros(36, dtype=np.int32) for k in KNOBS}
    P = C.POINTER(C.c_int)
    L.sqjk_get_tuning(eng._h, *[t[k].ctypes.data_as(P) for k in KNOBS])
    return {k: t[k].reshape(6, 6) for k in KNOBS}


def batch_runner(batch):
    def run(tables):
        batch.set_tuning(tables["rows"], tables["split"], tables["grid"])
        batch.build(graph=False)
        return batch.stats()["class_ms"]
    return run


def engine_runner(eng, P, cut):
    import ctypes as C
    from .batch import _bind
    L = _bind()
    P32 = C.POINTER(C.c_int)

    def run(tables):
        arr = [np.ascontiguousarray(tables[k].reshape(36), dtype=np.int32) for k in KNOBS]
        rc = L.sqjk_set_tuning(eng._h, *[a.ctypes.data_as(P32) for a in arr])
        if rc:
            raise RuntimeError("sqjk_set_tuning failed")
        eng.build(P, P, cut, cut, True)
        return eng.class_times()
    return run


def make_batch(count=256, seed=4256, cut=1e-10):
    """The batched workload: contexts, batch and seeded densities."""
    from . import inputs
    from .batch import JKBatch
    from .engine import JKEngine
    inst = inputs.batch_instances(count, seed)
    engines, bases = [], []
    for k, x in enumerate(inst):
        bas = inputs.make_basis(x["mol"], x["basis"])
        bases.append(bas)
        engines.append(JKEngine(bas.ptr, bas.n))
    batch = JKBatch(engines, [x["restricted"] for x in inst])
    dens = []
    for k, (x, e) in enumerate(zip(inst, engines)):
        rng = np.random.default_rng([seed, k])
        A = rng.standard_normal((e.n, e.n)) * 0.05
        B = rng.standard_normal((e.n, e.n)) * 0.05
        PA, PB = A + A.T, (B + B.T if not x["restricted"] else A + A.T)
        batch.set_density(k, PA, PB)
        dens.append((PA, PB))
    batch.set_cutoffs(cut, cut)
    return inst, bases, engines, batch, dens


def scf_density(mol, bas, conv=1e-6):
    """Alpha density of an RHF solve through the uhf() drop-in (the density
    the headline J/K build is timed on)."""
    import ctypes as C
    from .abi import GTOBasis, Molecule, Option, default_option, dptr
    from .engine import lib
    L = lib()
    L.uhf.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int, C.c_int] + \
        [C.POINTER(C.c_double)] * 4 + [C.POINTER(Option)]
    L.uhf.restype = C.c_double
    n, ne = bas.n, mol.n_electrons
    CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
    opt = default_option(SCFConv=conv)
    L.uhf(n, bas.ptr, mol.ptr, ne // 2, ne // 2, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt))
    occ = CA[:ne // 2]
    return occ.T @ occ


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["batch", "c60"])
    ap.add_argument("--out", default=None)
    ap.add_argument("--reps", type=int, default=2)
    a = ap.parse_args()
    t0 = time.time()
    if a.workload == "batch":
        inst, bases, engines, batch, _ = make_batch()
        batch.set_profile(True)
        batch.build(graph=False)
        start = _engine_tables(engines[0])
        tables, rep = tune(batch_runner(batch), start, a.reps)
        desc = "256 seeded small molecules (sqamd.inputs.batch_instances, seed 4256), cut-off 1e-10"
        out = a.out or os.path.join(DATA, "batch_c4.json")
    else:
        from . import inputs
        from .engine import JKEngine
        Z, X = inputs.c60()
        mol = inputs.mol_from_angstrom(Z, X)
        bas = inputs.make_basis(mol, "631g_star")
        P = scf_density(mol, bas)
        eng = JKEngine(bas.ptr, bas.n)
        eng.set_profile(True)
        start = _engine_tables(eng)
        tables, rep = tune(engine_runner(eng, P, 1e-10), start, a.reps)
        desc = "C60 RHF/6-31G*, 900 functions, converged RHF density (SCFConv 1e-6), cut-off 1e-10"
        out = a.out or os.path.join(DATA, "c60.json")
    save(out, tables, rep, desc)
    print(f"tuned in {time.time() - t0:.1f} s: {rep['before_ms']:.2f} -> {rep['after_ms']:.2f} ms; wrote {out}",
          flush=True)


if __name__ == "__main__":
    sys.exit(main())
