#!/usr/bin/env python3
"""J/K builds per second of the MI355X two-electron engine.

Workload (one "step" = one full J/K build, inputs resident on the host as in
the reference's uhf_getGMatrix call): seeded water cluster (H2O)32
(seed 3201, rigid monomers, O-O >= 2.7 A, 1 g/cm^3), RHF/6-31G* with
Cartesian d, 608 basis functions; Schwarz and density-weighted screening at
1e-12 (the reference's rules); seeded random symmetric density; FP64.

Prints one JSON line.  `roofline` is the FP64 vector-ALU roofline of the
integral kernels (modelled flops per build / HIP-event kernel time), and
`cpu_baseline` times the reference's own code (oracle/_ref, built from the
reference sources) on a bounded slice of the same build on one host core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "siam-quantum_amd"))

METRIC = "J/K builds/s & SCF solves/s (p50/p99 lat), mixed vs FP64 GPU at |ΔE|≤1µEh"
PEAK_FP64_TFLOPS = 78.6
WORKLOADS = {
    "w32": (32, 3201),
    "w16": (16, 1601),
}


def make_inputs(name):
    from sqamd import inputs
    nmol, seed = WORKLOADS[name]
    Z, X = inputs.water_cluster(nmol, seed)
    mol = inputs.mol_from_angstrom(Z, X)
    bas = inputs.make_basis(mol, "631g_star")
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((bas.n, bas.n)) * 0.01
    return mol, bas, A + A.T


def scf_solves(name, nsolve, mixed_thr):
    """Repeated RHF SCF solves through the uhf() drop-in (reference defaults:
    3-step cut-offs, DIIS4, SCFConv 1e-8): each solve builds its per-molecule
    state, runs the SCF on the device and releases the state, as uhf() does.
    Returns FP64 and mixed-precision solves/s with p50/p99 latency and the
    energy difference."""
    import ctypes as C
    from sqamd import inputs, lib, default_option, dptr, GTOBasis, Molecule, Option
    nmol, seed = WORKLOADS[name]
    Z, X = inputs.water_cluster(nmol, seed)
    mol = inputs.mol_from_angstrom(Z, X)
    bas = inputs.make_basis(mol, "631g_star")
    L = lib()
    L.uhf.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int, C.c_int] + \
        [C.POINTER(C.c_double)] * 4 + [C.POINTER(Option)]
    L.uhf.restype = C.c_double
    L.sqjk_set_dropin_mixed.argtypes = [C.c_double]
    n, ne = bas.n, mol.n_electrons
    out = {"workload": f"(H2O){nmol} RHF/6-31G* SCF, {n} basis functions, seed {seed}, "
                       f"reference defaults with SCFConv 1e-8", "solves": nsolve}
    for label, thr in (("fp64", 0.0), ("mixed", mixed_thr)):
        if label == "mixed" and thr <= 0:
            continue
        L.sqjk_set_dropin_mixed(thr)
        lat, E = [], None
        for _ in range(nsolve):
            CA, CB, eA, eB = np.zeros((n, n)), np.zeros((n, n)), np.zeros(n), np.zeros(n)
            opt = default_option(SCFConv=1e-8)
            t0 = time.perf_counter()
            E = L.uhf(n, bas.ptr, mol.ptr, ne // 2, ne // 2, dptr(CA), dptr(CB), dptr(eA), dptr(eB), C.byref(opt))
            lat.append(time.perf_counter() - t0)
        lat = np.array(lat)
        out[label] = {"value": len(lat) / lat.sum(), "unit": "SCF solves/s",
                      "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                      "energy": E}
        if label == "mixed":
            out[label]["threshold"] = thr
            out[label]["abs_dE"] = abs(E - out["fp64"]["energy"])
    L.sqjk_set_dropin_mixed(0.0)
    return out


def cpu_baseline(bas, P, cut, slices):
    """Reference code (oracle/_ref) on the PQ % slices == 0 part of the same
    build, one core; builds/s scaled by the number of slices."""
    import ctypes as C
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
    return {"value": 1.0 / (dt * slices), "unit": "builds/s", "cores": 1, "kind": "reference",
            "sample": f"reference GTO_JK_Matrix_Quartet_Parallel, childID 0 of nCPU={slices} "
                      f"(bra shell pairs PQ % {slices} == 0) of the same build, {dt:.1f} s, x{slices}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="w32", choices=sorted(WORKLOADS))
    ap.add_argument("--cutoff", type=float, default=1e-12)
    ap.add_argument("--cpu-slices", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed-threshold", type=float, default=1e-4,
                    help="Schwarz product below which quartets run in FP32 for the mixed line (0: skip)")
    ap.add_argument("--scf-solves", type=int, default=5, help="SCF solves per precision (0: skip)")
    ap.add_argument("--scf-workload", default="w16", choices=sorted(WORKLOADS))
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
    mol, bas, P = make_inputs(a.workload)
    n = bas.n
    eng = JKEngine(bas.ptr, n)
    eng.set_profile(True)
    for _ in range(a.warmup):
        eng.build(P, P, a.cutoff, a.cutoff, True)
    kms, cms = [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        eng.build(P, P, a.cutoff, a.cutoff, True)
        kms.append(eng.stats()["kernel_ms"])
        cms.append(eng.class_times())
    wall = time.perf_counter() - t0
    st = eng.stats()
    cls = eng.class_stats()
    flops = workmodel.build_flops(cls)
    kernel_s = float(np.mean(kms)) / 1e3
    achieved = flops / kernel_s / 1e12
    # dominant class kernel: HIP events on the engine's stream around its launches
    cls_ms = np.mean(np.array(cms), axis=0)
    dom = int(np.argmax(cls_ms))
    dlq, dnq, dnpq = cls[dom]
    dflops = dnpq * workmodel.flops_per_prim_quartet(*dlq) + dnq * workmodel.digestion_flops(*dlq)
    dach = dflops / (cls_ms[dom] / 1e3) / 1e12
    ms = wall / a.steps * 1e3
    out = {
        "metric": METRIC,
        "value": a.steps / wall,
        "unit": "J/K builds/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"(H2O){WORKLOADS[a.workload][0]} RHF/6-31G* full J/K build, "
                               f"{n} basis functions, seed {WORKLOADS[a.workload][1]}, "
                               f"Schwarz+density screening {a.cutoff:g}, FP64",
                   "basis_functions": n, "quartets": st["quartets"],
                   "integrals": workmodel.integrals(cls)},
        "classes": {"(%d%d|%d%d)" % lq: {"quartets": nq, "prim_quartets": npq,
                                         "gflop": (npq * workmodel.flops_per_prim_quartet(*lq) +
                                                   nq * workmodel.digestion_flops(*lq)) / 1e9,
                                         "kernel_ms": float(cls_ms[k])}
                    for k, (lq, nq, npq) in enumerate(cls)},
        "kernel_ms": kernel_s * 1e3,
        "roofline": {"bound": "valu-fp64", "kernel": "k_class (%d%d|%d%d) FP64" % dlq,
                     "achieved": dach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": dach / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel_ms": float(cls_ms[dom]), "flops_per_launch": dflops,
                     "build": {"achieved": achieved, "frac": achieved / PEAK_FP64_TFLOPS,
                               "flops_per_build": flops, "kernel_ms": kernel_s * 1e3}},
    }
    if a.mixed_threshold > 0:
        eng.set_mixed(a.mixed_threshold)
        eng.build(P, P, a.cutoff, a.cutoff, True)
        J64 = eng.build(P, P, a.cutoff, a.cutoff, True)  # warm, then compare below
        t1 = time.perf_counter()
        mk = []
        for _ in range(a.steps):
            Jm, Km, _ = eng.build(P, P, a.cutoff, a.cutoff, True)
            mk.append(eng.stats()["kernel_ms"])
        wall_m = time.perf_counter() - t1
        stm = eng.stats()
        eng.set_mixed(0.0)
        Jd, Kd, _ = eng.build(P, P, a.cutoff, a.cutoff, True)
        out["mixed"] = {"value": a.steps / wall_m, "unit": "J/K builds/s", "threshold": a.mixed_threshold,
                        "kernel_ms": float(np.mean(mk)), "quartets_fp32": stm["quartets_fp32"],
                        "max_abs_dJ": float(np.abs(Jm - Jd).max()), "max_abs_dK": float(np.abs(Km - Kd).max()),
                        "speedup_vs_fp64": (a.steps / wall_m) / out["value"]}
    eng.close()
    if a.scf_solves > 0:
        out["scf"] = scf_solves(a.scf_workload, a.scf_solves, a.mixed_threshold)
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(bas, P, a.cutoff, a.cpu_slices)
        except FileNotFoundError as e:
            out["cpu_baseline"] = {"value": None, "unit": "builds/s", "cores": 1, "kind": "reference",
                                   "sample": f"not run: {e}"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
