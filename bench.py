#!/usr/bin/env python3
"""J/K builds per second of the MI355X two-electron engine.

Headline workload (one "step" = one full J/K build): the C60 truncated
icosahedron (C-C 1.45/1.40 A), RHF/6-31G* with Cartesian d, 900 basis
functions, Schwarz and density-weighted screening at 1e-10 (the reference's
rules), FP64, on the density of a converged SCF of the same molecule
(computed first by the uhf() drop-in on the device).  Inputs are host
arrays, as in the reference's uhf_getGMatrix call, so `value` includes the
density upload and the J/K download.

Prints one JSON line.  `roofline` is the FP64 vector-ALU roofline of the
dominant class kernel: the useful work of one launch (surviving primitive
quartets x modelled flops per primitive quartet + surviving quartets x
digestion flops) over its HIP-event time on the engine's stream;
`traffic` is the measured HBM bytes of that launch from the committed
rocprofv3 counter summary.  `cpu_baseline` times the reference's own code
(oracle/_ref, built from the reference sources) on a bounded slice of the
same build on one host core.
Further keys: the (H2O)32 build, the incremental (Delta-D) build, the mixed
FP32/FP64 build, SCF solves/s with p50/p99 latency.
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "siam-quantum_amd"))

METRIC = "J/K builds/s & SCF solves/s (p50/p99 lat), mixed vs FP64 GPU at |ΔE|≤1µEh"
PEAK_FP64_TFLOPS = 78.6
PROFILES = os.path.join(ROOT, "profiles")


def molecule(name):
    from sqamd import inputs
    if name == "c60":
        Z, X = inputs.c60()
        desc = "C60 truncated icosahedron (C-C 1.45/1.40 A)"
    else:
        nm, seed = {"w16": (16, 1601), "w32": (32, 3201)}[name]
        Z, X = inputs.water_cluster(nm, seed)
        desc = f"(H2O){nm} cluster, seed {seed}"
    mol = inputs.mol_from_angstrom(Z, X)
    return mol, inputs.make_basis(mol, "631g_star"), desc


def uhf_fn():
    from sqamd import lib, GTOBasis, Molecule, Option
    L = lib()
    L.uhf.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int, C.c_int] + \
        [C.POINTER(C.c_double)] * 4 + [C.POINTER(Option)]
    L.uhf.restype = C.c_double
    L.sqjk_set_dropin_mixed.argtypes = [C.c_double]
    return L


def scf_once(mol, bas, orbitals=None, **kw):
    """RHF through the uhf() drop-in; returns (E, density P_alpha, seconds).
    With a dict `orbitals`, the orbitals, orbital energies and options are
    stored in it (CA, CB, eA, eB, opt)."""
    from sqamd import default_option, dptr
    L = uhf_fn()
    n, ne = bas.n, mol.n_electrons
    CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
    opt = default_option(**kw)
    t0 = time.perf_counter()
    E = L.uhf(n, bas.ptr, mol.ptr, ne // 2, ne // 2, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt))
    dt = time.perf_counter() - t0
    if orbitals is not None:
        orbitals.update(CA=CA, CB=CB, eA=eA, eB=eB, opt=opt)
    occ = CA[:ne // 2]
    return E, occ.T @ occ, dt


def mp2_bench(name, steps):
    """Frozen-core RHF MP2 energy on the device (sqjk_mp2_rhf, Schwarz cut-off
    1e-10) on the orbitals of a converged SCF (uhf() drop-in, SCFConv 1e-8):
    (H2O)16/6-31G*, or C60/3-21G, the molecule of the reference's published
    MP2 timing (47 min on two i7 cores, docs/FAQ.md:121-129); "w16_cation":
    unrestricted MP2 (sqjk_mp2_uhf) of the (H2O)16 cation doublet on its
    UHF orbitals."""
    from sqamd import JKEngine, inputs
    from sqamd.engine import frozen_core
    uhf = name == "w16_cation"
    if uhf:
        mol0, _, d = molecule("w16")
        mol = inputs.mol_from_angstrom(mol0.Z, np.stack([mol0.x, mol0.y, mol0.z], 1) / inputs.ANGSTROM2BOHR, 1)
        bas = inputs.make_basis(mol, "631g_star")
        desc = f"{d} cation doublet UHF/6-31G*"
    elif name == "c60_321g":
        Z, X = inputs.c60()
        mol = inputs.mol_from_angstrom(Z, X)
        bas = inputs.make_basis(mol, "321g")
        desc = "C60 truncated icosahedron (C-C 1.45/1.40 A) RHF/3-21G"
    else:
        mol, _, d = molecule(name)
        bas = inputs.make_basis(mol, "631g_star")
        desc = f"{d} RHF/6-31G*"
    orbs = {}
    ne, ncore = mol.n_electrons, frozen_core(mol)
    if uhf:
        from sqamd import default_option, dptr
        n = bas.n
        nea, neb = (ne + 1) // 2, ne // 2
        CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
        opt = default_option(SCFConv=1e-8, restricted=0, multiplicity=2)
        t0 = time.perf_counter()
        E = uhf_fn().uhf(n, bas.ptr, mol.ptr, nea, neb, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt))
        t_scf = time.perf_counter() - t0
        nocc = nea
    else:
        E, _, t_scf = scf_once(mol, bas, orbs, SCFConv=1e-8)
        nocc = ne // 2
    eng = JKEngine(bas.ptr, bas.n)
    runs = []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        try:
            if uhf:
                r = eng.mp2_uhf(nea, neb, ncore, eA, eB, CA, CB, 1e-10)
            else:
                r = eng.mp2_rhf(nocc, ncore, orbs["eA"], orbs["CA"], 1e-10)
        except RuntimeError as e:
            eng.close()
            return {"workload": desc, "error": str(e)}
        r["seconds"] = time.perf_counter() - t0
        runs.append(r)
    eng.close()
    r = min(runs, key=lambda x: x["seconds"])
    out = {"workload": f"{desc} frozen-core MP2 on the converged SCF, FP64, Schwarz 1e-10",
           "basis_functions": bas.n, "correlated_occupied": nocc - ncore, "virtual": bas.n - nocc,
           **({"correlated_occupied_beta": neb - ncore, "virtual_beta": bas.n - neb} if uhf else {}),
           "scf_energy": E, "scf_seconds": t_scf, "mp2_energy": r["e2"], "seconds": r["seconds"],
           "runs": len(runs), "integral_store_ms": r["store_ms"], "first_half_ms": r["half1_ms"],
           "second_half_ms": r["half2_ms"], "second_half_tflops": r["half2_flop"] / (r["half2_ms"] * 1e-3) / 1e12,
           "first_half_tflops": r["half1_flop"] / (r["half1_ms"] * 1e-3) / 1e12,
           "quartets": r["quartets"], "device_gb": r["device_bytes"] / 1e9}
    if name == "c60_321g":
        out["reference_published"] = "47 min, two i7 @3.4 GHz cores, 6 GB (docs/FAQ.md:121-129)"
    return out


def forces_bench(eng, mol, bas, P, orbs, steps):
    """Hartree-Fock forces of the converged C60 SCF: the two-electron
    gradient (sqjk_grad_eri on the build context, cut-off 1e-10 as the
    reference's GRAD_SCHWARZ_CUTOFF), the one-electron gradient
    (sqjk_grad_one_electron) and the whole uhf_force drop-in (host densities,
    energy-weighted density, both device parts, host arrays in and out)."""
    from sqamd.engine import basis_to_atom, grad_one_electron, uhf_force
    n, na = bas.n, mol.ptr.contents.nAtom
    b2a = basis_to_atom(mol, bas)
    nocc = mol.n_electrons // 2
    CA, eA = orbs["CA"], orbs["eA"]
    QT = 2.0 * (CA[:nocc].T * eA[:nocc]) @ CA[:nocc]
    eng.set_profile(False)
    eng.grad_eri(P, P, b2a, na, 1.0, 1e-10)  # warm-up (plan for the gradient cut-off)
    t2, k2 = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        G2 = eng.grad_eri(P, P, b2a, na, 1.0, 1e-10)
        t2.append(time.perf_counter() - t0)
        k2.append(eng.stats()["kernel_ms"])
    grad_one_electron(bas, mol, 2 * P, QT, b2a)
    t1 = []
    for _ in range(steps):
        t0 = time.perf_counter()
        G1 = grad_one_electron(bas, mol, 2 * P, QT, b2a)
        t1.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    F = uhf_force(bas, mol, nocc, nocc, orbs["CA"], orbs["CB"], orbs["eA"], orbs["eB"], orbs["opt"])
    tf = time.perf_counter() - t0
    return {"workload": "C60 RHF/6-31G* forces on the converged SCF (SCFConv 1e-6), FP64",
            "two_electron_ms": 1e3 * float(np.median(t2)), "two_electron_kernel_ms": float(np.median(k2)),
            "one_electron_ms": 1e3 * float(np.median(t1)), "uhf_force_ms": 1e3 * tf,
            "max_abs_force": float(np.abs(F).max()), "sum_forces": float(np.abs(F.sum(0)).max()),
            "max_abs_grad_two_electron": float(np.abs(G2).max()), "max_abs_grad_one_electron": float(np.abs(G1).max())}


def scf_solves(name, nsolve, mixed_thr):
    """Repeated RHF SCF solves through the uhf() drop-in (reference defaults:
    3-step cut-offs, DIIS4, SCFConv 1e-8).  The first (cold) solve is
    reported apart; p50/p99 are over the warm ones."""
    from sqamd import default_option, dptr
    mol, bas, desc = molecule(name)
    L = uhf_fn()
    n, ne = bas.n, mol.n_electrons
    out = {"workload": f"{desc} RHF/6-31G* SCF, {n} basis functions, reference defaults with SCFConv 1e-8",
           "warm_solves": nsolve}
    for label, thr in (("fp64", 0.0), ("mixed", mixed_thr)):
        if label == "mixed" and thr <= 0:
            continue
        L.sqjk_set_dropin_mixed(thr)
        lat, E = [], None
        for _ in range(nsolve + 1):
            CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
            opt = default_option(SCFConv=1e-8)
            t0 = time.perf_counter()
            E = L.uhf(n, bas.ptr, mol.ptr, ne // 2, ne // 2, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt))
            lat.append(time.perf_counter() - t0)
        cold, warm = lat[0], np.array(lat[1:])
        out[label] = {"value": len(warm) / warm.sum(), "unit": "SCF solves/s",
                      "p50_ms": float(np.percentile(warm, 50) * 1e3), "p99_ms": float(np.percentile(warm, 99) * 1e3),
                      "cold_ms": cold * 1e3, "energy": E}
        if label == "mixed":
            out[label]["threshold"] = thr
            out[label]["abs_dE"] = abs(E - out["fp64"]["energy"])
    L.sqjk_set_dropin_mixed(0.0)
    return out


def profile_traffic(kernel):
    """HBM-side bytes of `kernel` per build (all of its launches in one
    build) from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE runs, calibrated against a plain copy), or None when no
    summary names this kernel."""
    path = os.path.join(PROFILES, "jk_c60_631gs_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        m = re.match(r"\s*traffic_bytes_per_launch\s*\|\s*(.+?)\s*\|\s*(\d+)\s*\|\s*(\d+)", line)
        if m and m.group(1) == kernel:
            return int(m.group(2)) + int(m.group(3)), os.path.relpath(path, ROOT)
    return None, None


def timed_builds(eng, P, cut, steps, warmup):
    for _ in range(warmup):
        eng.build(P, P, cut, cut, True)
    kms, cms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.build(P, P, cut, cut, True)
        kms.append(eng.stats()["kernel_ms"])
        cms.append(eng.class_times())
    return time.perf_counter() - t0, float(np.mean(kms)), np.mean(np.array(cms), axis=0)


def class_label(lq):
    return "(%d%d|%d%d)" % tuple(lq)


def cpu_baseline(bas, P, cut, slices):
    """Reference code (oracle/_ref) on the PQ % slices == 0 part of the same
    build, one core; builds/s scaled by the number of slices."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Reference, default_option
    from sqamd import dptr
    ref = Reference()
    n = bas.n
    sch = ref.schwarz(bas.ptr, n)
    opt = default_option(SCFCutoff=cut, restricted=1, maxMem=1, nCPU=slices)
    ref.jk_reset(opt)
    J, K1, K2 = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    t0 = time.perf_counter()
    ref.L.GTO_JK_Matrix_Quartet_Parallel(0, n, dptr(P), dptr(P), bas.ptr, dptr(sch), cut,
                                         dptr(J), dptr(K1), dptr(K2), C.byref(opt))
    dt = time.perf_counter() - t0
    ref.jk_reset(opt)
    return {"value": 1.0 / (dt * slices), "unit": "J/K builds/s", "cores": 1, "kind": "reference",
            "host_cores_visible": os.cpu_count(),
            "sample": f"reference GTO_JK_Matrix_Quartet_Parallel, childID 0 of nCPU={slices} "
                      f"(bra shell pairs PQ % {slices} == 0) of the same C60 build, {dt:.1f} s on one core, "
                      f"x{slices}"}


def batch_bench(steps, mixed_thr, cut=1e-10):
    """Config C4: 256 seeded small molecules (a quarter UHF doublet cations)
    as one graph-captured set of class launches, against the same builds
    one molecule at a time through the plain path (sqjk_build, FP64, host
    arrays in and out, as uhf_getGMatrix does)."""
    from sqamd import tune
    inst, bases, engines, batch, dens = tune.make_batch(cut=cut)
    out = {"workload": f"{len(inst)} seeded molecules (H2O..C6H6, STO-3G/6-31G/6-31G*, "
                       f"{min(b.n for b in bases)}-{max(b.n for b in bases)} functions, "
                       f"{sum(not x['restricted'] for x in inst)} UHF doublet cations, sigma 0.05 A), "
                       f"seeded densities, cut-off {cut:g}",
           "molecules": len(inst), "basis_functions": int(sum(b.n for b in bases))}
    path = os.path.join(tune.DATA, "batch_c4.json")
    if os.path.exists(path):
        t = tune.load(path)
        batch.set_tuning(t["rows"], t["split"], t["grid"])
        out["tuning"] = os.path.relpath(path, ROOT)
    # the plain FP64 path, one molecule at a time
    ref = []
    for e, (PA, PB), x in zip(engines, dens, inst):
        ref.append(e.build(PA, PB, cut, cut, x["restricted"]))
    t0 = time.perf_counter()
    for e, (PA, PB), x in zip(engines, dens, inst):
        e.build(PA, PB, cut, cut, x["restricted"])
    t_seq = time.perf_counter() - t0
    out["per_molecule_fp64"] = {"seconds": t_seq, "molecule_builds_per_s": len(inst) / t_seq}
    for label, thr in (("batched_fp64", 0.0), ("batched_mixed", mixed_thr)):
        if label == "batched_mixed" and thr <= 0:
            continue
        batch.set_mixed(thr)
        batch.build(graph=True)
        batch.build(graph=True)
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.build(graph=True)
            kms.append(batch.stats()["kernel_ms"])
        dt = (time.perf_counter() - t0) / steps
        d = 0.0
        for k in range(0, len(inst), 17):
            J, KA, KB = batch.result(k)
            d = max(d, float(np.abs(J - ref[k][0]).max()), float(np.abs(KA - ref[k][1]).max()),
                    float(np.abs(KB - ref[k][2]).max()))
        out[label] = {"ms_per_batch": dt * 1e3, "kernel_ms": float(np.mean(kms)),
                      "molecule_builds_per_s": len(inst) / dt, "speedup_vs_per_molecule_fp64": t_seq / dt,
                      "max_abs_diff_vs_per_molecule_fp64": d, "threshold": thr,
                      "launches": batch.stats()["launches"]}
    batch.set_mixed(0.0)
    batch.close()
    for e in engines:
        e.close()
    return out


def serve_bench(count, rate, mixed_thr, slots):
    """Config C5: Poisson arrivals of RHF/UHF 6-31G requests (random C/H/N/O
    chains, 10-200 atoms log-uniform, two geometry steps each), served with
    device-resident SCF state, admission by device-memory footprint.  The same
    arrival trace runs in FP64 and in mixed precision."""
    from sqamd import inputs
    from sqamd.serve import serve
    reqs = inputs.serving_requests(count=count, rate=rate, seed=5005, steps=2)
    out = {"workload": f"{count} requests, Poisson arrivals at {rate:g}/s (seed 5005), random C/H/N/O chains "
                       f"of {min(r['natom'] for r in reqs)}-{max(r['natom'] for r in reqs)} atoms, 6-31G, "
                       f"2 geometry steps (sigma 0.05 A) each, SCFConv 1e-6, 3-step cut-offs, DIIS4",
           "slots": slots, "uhf_requests": int(sum(r["multiplicity"] == 2 for r in reqs))}
    E = {}
    for label, thr in (("fp64", 0.0), ("mixed", mixed_thr)):
        if label == "mixed" and thr <= 0:
            continue
        t0 = time.perf_counter()
        res = serve(reqs, slots=slots, mixed=thr, conv=1e-6, mem_budget=96e9)
        wall = time.perf_counter() - t0
        lat = np.array([st["finish"] - st["arrival"] for _, _, st in res])
        solves = sum(len(e) for e, _, _ in res)
        conv = sum(st["converged_steps"] for _, _, st in res)
        makespan = max(st["finish"] for _, _, st in res)
        E[label] = np.concatenate([e for e, _, _ in res])
        out[label] = {"scf_solves_per_s": solves / makespan, "requests_per_s": len(res) / makespan,
                      "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                      "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                      "converged_steps": int(conv), "steps": int(solves), "makespan_s": makespan, "wall_s": wall,
                      "peak_in_flight": res[0][2]["peak_in_flight"],
                      "peak_memory_GB": res[0][2]["peak_memory"] / 1e9,
                      "basis_functions": [int(st["nBasis"]) for _, _, st in res],
                      "iterations": int(sum(st["iterations"] for _, _, st in res))}
    if "mixed" in E:
        ok = (E["fp64"] != 0) & (E["mixed"] != 0)
        out["mixed"]["max_abs_dE_vs_fp64"] = float(np.abs(E["mixed"][ok] - E["fp64"][ok]).max()) if ok.any() else None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cutoff", type=float, default=1e-10)
    ap.add_argument("--cpu-slices", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed-threshold", type=float, default=1e-4,
                    help="Schwarz product below which quartets run in FP32 for the mixed line (0: skip)")
    ap.add_argument("--scf-solves", type=int, default=100, help="warm SCF solves per precision (0: skip)")
    ap.add_argument("--scf-workload", default="w16", choices=["w16", "w32"])
    ap.add_argument("--no-extras", action="store_true", help="only the headline C60 build line")
    ap.add_argument("--no-tuning", action="store_true", help="default work-plan tables, not the tuned ones")
    ap.add_argument("--serve-requests", type=int, default=16)
    ap.add_argument("--serve-rate", type=float, default=4.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--forces-only", action="store_true",
                    help="only the C60 SCF and its forces (a JSON line with the forces timings), for profiling")
    ap.add_argument("--mp2-only", choices=["w16", "w16_cation", "c60_321g"], default=None,
                    help="only the MP2 benchmark of this workload")
    ap.add_argument("--no-mp2-c60", action="store_true", help="leave out the C60/3-21G MP2 extra")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 times), to calibrate "
                         "the rocprofv3 FETCH_SIZE/WRITE_SIZE counters, then exit")
    a = ap.parse_args()
    if a.copy_calibration_mib:
        import torch
        nbytes = a.copy_calibration_mib << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 5}))
        return
    if a.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    from sqamd import JKEngine, workmodel
    if a.mp2_only:
        print(json.dumps({"mp2": mp2_bench(a.mp2_only, a.steps)}))
        return
    mol, bas, desc = molecule("c60")
    n = bas.n
    print(f"bench: C60 SCF for the step density (n={n})", file=sys.stderr, flush=True)
    orbs = {}
    E_scf, P, t_scf = scf_once(mol, bas, orbs, SCFConv=1e-6)
    if a.forces_only:
        eng = JKEngine(bas.ptr, n)
        from sqamd import tune
        tpath = os.path.join(tune.DATA, "c60.json")
        if os.path.exists(tpath) and not a.no_tuning:
            t = tune.load(tpath)
            eng.set_tuning(t["rows"], t["split"], t["grid"])
        print(json.dumps({"forces": forces_bench(eng, mol, bas, P, orbs, a.steps), "scf_energy": E_scf}))
        eng.close()
        return
    print(f"bench: C60 E = {E_scf:.8f} Eh in {t_scf:.1f} s; timing builds", file=sys.stderr, flush=True)
    eng = JKEngine(bas.ptr, n)
    eng.set_profile(True)
    tuning = None
    from sqamd import tune
    tpath = os.path.join(tune.DATA, "c60.json")
    if os.path.exists(tpath) and not a.no_tuning:
        t = tune.load(tpath)
        eng.set_tuning(t["rows"], t["split"], t["grid"])
        tuning = os.path.relpath(tpath, ROOT)
    wall, kernel_ms, cls_ms = timed_builds(eng, P, a.cutoff, a.steps, a.warmup)
    st = eng.stats()
    plan = eng.class_stats()
    work = eng.class_work()
    # useful work: surviving primitive quartets and quartets of each class
    cflops = [wp * workmodel.flops_per_prim_quartet(*lq) + wq * workmodel.digestion_flops(*lq)
              for lq, wq, wp, _, _ in work]
    dom = int(np.argmax(cls_ms))
    dlq = work[dom][0]
    if dlq == (2, 2, 2, 2):
        dname = "k_class_wave<2, 2, 2, 2, 0, double>" if os.environ.get("SQJK_DD_KERNEL") == "wave" \
            else "k_class_wd<2, 2, 2, 2, 0>"
    else:
        dname = "k_class<%d, %d, %d, %d, 0, double, %d>" % (dlq + (512 if workmodel.class_block(*dlq) == 512 else 256,))
    dach = cflops[dom] / (cls_ms[dom] / 1e3) / 1e12
    build_ach = sum(cflops) / (kernel_ms / 1e3) / 1e12
    traffic, traffic_src = profile_traffic(dname)
    out = {
        "metric": METRIC,
        "value": a.steps / wall,
        "unit": "J/K builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{desc} RHF/6-31G* full J/K build, {n} basis functions, Schwarz+density "
                               f"screening {a.cutoff:g}, converged-SCF density, FP64, host arrays in/out",
                   "basis_functions": n, "quartets_screened": st["quartets"],
                   "quartets_surviving": int(sum(w[1] for w in work)),
                   "prim_quartets_surviving": int(sum(w[2] for w in work)),
                   "integrals_surviving": int(sum(wq * workmodel.nint(*lq) for lq, wq, _, _, _ in work)),
                   "scf_energy": E_scf, "scf_seconds": t_scf, "tuning": tuning},
        "kernel_ms": kernel_ms,
        "classes": {class_label(lq): {"quartets_plan": plan[k][1], "quartets": wq, "prim_quartets": wp,
                                      "gflop": cflops[k] / 1e9, "kernel_ms": float(cls_ms[k]),
                                      "frac": cflops[k] / (cls_ms[k] / 1e3) / 1e12 / PEAK_FP64_TFLOPS
                                      if cls_ms[k] > 0 else None}
                    for k, (lq, wq, wp, _, _) in enumerate(work)},
        "roofline": {"bound": "valu-fp64", "kernel": dname,
                     "achieved": dach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": dach / PEAK_FP64_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel_ms": float(cls_ms[dom]), "flops_per_launch": cflops[dom],
                     "units_per_launch": {"quartets": work[dom][1], "prim_quartets": work[dom][2]},
                     "build": {"achieved": build_ach, "frac": build_ach / PEAK_FP64_TFLOPS,
                               "flops_per_build": sum(cflops), "kernel_ms": kernel_ms}},
    }
    if not a.no_extras:
        # incremental build: a late-iteration density change (1e-4 scale)
        rng = np.random.default_rng(60)
        A = rng.standard_normal((n, n)) * 1e-4
        dP = A + A.T
        timed_builds(eng, dP, a.cutoff, 1, 1)
        t0 = time.perf_counter()
        for _ in range(a.steps):
            eng.build(dP, dP, a.cutoff, a.cutoff, True)
        out["delta_d"] = {"ms_per_build": (time.perf_counter() - t0) / a.steps * 1e3,
                          "kernel_ms": eng.stats()["kernel_ms"],
                          "quartets_surviving": int(sum(w[1] for w in eng.class_work())),
                          "density": "P_conv-like change: seeded symmetric N(0, 1e-4)"}
        if a.mixed_threshold > 0:
            J64, K64, _ = eng.build(P, P, a.cutoff, a.cutoff, True)
            eng.set_mixed(a.mixed_threshold)
            eng.build(P, P, a.cutoff, a.cutoff, True)
            t1 = time.perf_counter()
            mk = []
            for _ in range(a.steps):
                Jm, Km, _ = eng.build(P, P, a.cutoff, a.cutoff, True)
                mk.append(eng.stats()["kernel_ms"])
            wall_m = time.perf_counter() - t1
            stm = eng.stats()
            eng.set_mixed(0.0)
            out["mixed"] = {"value": a.steps / wall_m, "unit": "J/K builds/s", "threshold": a.mixed_threshold,
                            "kernel_ms": float(np.mean(mk)), "quartets_fp32": stm["quartets_fp32"],
                            "max_abs_dJ": float(np.abs(Jm - J64).max()), "max_abs_dK": float(np.abs(Km - K64).max()),
                            "speedup_vs_fp64": (a.steps / wall_m) / out["value"]}
    if not a.no_extras:
        print("bench: C60 forces", file=sys.stderr, flush=True)
        out["forces"] = forces_bench(eng, mol, bas, P, orbs, a.steps)
    eng.close()
    if not a.no_extras:
        mol32, bas32, desc32 = molecule("w32")
        rng = np.random.default_rng(3201)
        A = rng.standard_normal((bas32.n, bas32.n)) * 0.01
        e32 = JKEngine(bas32.ptr, bas32.n)
        w32, k32, _ = timed_builds(e32, A + A.T, 1e-12, a.steps, a.warmup)
        e32.close()
        out["w32"] = {"workload": f"{desc32} RHF/6-31G*, {bas32.n} functions, screening 1e-12, seeded density",
                      "value": a.steps / w32, "unit": "J/K builds/s", "kernel_ms": k32}
        print("bench: batched builds", file=sys.stderr, flush=True)
        out["batch"] = batch_bench(a.steps, a.mixed_threshold)
        print("bench: serving", file=sys.stderr, flush=True)
        out["serve"] = serve_bench(a.serve_requests, a.serve_rate, a.mixed_threshold, a.serve_slots)
        if a.scf_solves > 0:
            print("bench: SCF solves", file=sys.stderr, flush=True)
            out["scf"] = scf_solves(a.scf_workload, a.scf_solves, a.mixed_threshold)
        print("bench: MP2 (H2O)16/6-31G*", file=sys.stderr, flush=True)
        out["mp2"] = mp2_bench("w16", 2)
        out["mp2_uhf"] = mp2_bench("w16_cation", 2)
        if not a.no_mp2_c60:
            print("bench: MP2 C60/3-21G", file=sys.stderr, flush=True)
            out["mp2_c60"] = mp2_bench("c60_321g", 1)
    if not a.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        try:
            out["cpu_baseline"] = cpu_baseline(bas, P, a.cutoff, a.cpu_slices)
        except FileNotFoundError as e:
            out["cpu_baseline"] = {"value": None, "unit": "J/K builds/s", "cores": 1, "kind": "reference",
                                   "sample": f"not run: {e}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
