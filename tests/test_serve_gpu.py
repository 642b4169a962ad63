"""The serving loop (sqsrv_*): SCF solves of requests that arrive over
time, admitted by slot and device-memory footprint, run side by side.

- Each request's energies are the same alone (one slot, arrivals spread
  out) and in the mix (four slots, all at once): 1e-9 Eh.  Requests in the
  mix share the device, not state.
- A memory budget below any request admits one request at a time and gives
  the same energies.
- Each step of a request equals the reference's geometry loop run through
  the uhf() drop-in with the same settings: the first step from the DIAG
  guess, later ones from the previous step's orbitals (SCFGUESS_CACB,
  src/qmd.c:585-589): 1e-8 Eh.  (A UHF radical can have several SCF
  solutions; a later step follows the state of the previous one, as in the
  reference, and need not reach the one a DIAG start finds.)  Later steps
  take no more iterations than the first.
- Mixed precision (FP32 below a Schwarz product of 1e-4): |dE| <= 1e-6 Eh
  against FP64 for every step of every request.
"""
import ctypes as C

import numpy as np
import pytest

from sqamd import GTOBasis, Molecule, Option, default_option, dptr, inputs, lib
from sqamd.abi import SCFGUESS_CACB, SCFGUESS_DIAG
from sqamd.serve import serve

pytestmark = pytest.mark.gpu
CONV = 1e-8
MAXIT = 300  # one seeded UHF doublet does not converge within 80 DIIS iterations from the DIAG guess


@pytest.fixture(scope="module")
def requests():
    return inputs.serving_requests(count=4, rate=100.0, seed=77, steps=2, atoms=(6, 12))


def _uhf_steps(Z, xyz_steps, mult):
    """The reference's geometry loop through the uhf() drop-in: the first
    step from the DIAG guess, every later one from the previous step's
    orbitals (SCFGUESS_CACB, src/qmd.c:585-589)."""
    L = lib()
    L.uhf.argtypes = [C.c_int, C.POINTER(GTOBasis), C.POINTER(Molecule), C.c_int, C.c_int] + \
        [C.POINTER(C.c_double)] * 4 + [C.POINTER(Option)]
    L.uhf.restype = C.c_double
    ne = int(np.sum(Z))
    out, CA, CB = [], None, None
    for k, xyz in enumerate(xyz_steps):
        mol = inputs.Mol(Z, xyz)
        bas = inputs.make_basis(mol, "631g")
        n = bas.n
        if CA is None:
            CA, CB = np.zeros((n, n)), np.zeros((n, n))
        eA, eB = np.zeros(n), np.zeros(n)
        opt = default_option(SCFConv=CONV, restricted=int(mult == 1), SCFMax=MAXIT,
                             SCFGuess=SCFGUESS_DIAG if k == 0 else SCFGUESS_CACB)
        out.append(L.uhf(n, bas.ptr, mol.ptr, (ne + mult - 1) // 2, (ne - mult + 1) // 2, dptr(CA), dptr(CB),
                         dptr(eA), dptr(eB), C.byref(opt)))
    return out


def test_alone_equals_mix(requests):
    alone = []
    for r in requests:
        one = dict(r, arrival=0.0)
        alone.append(serve([one], slots=1, conv=CONV, max_iter=MAXIT)[0])
    mix = serve([dict(r, arrival=0.0) for r in requests], slots=4, conv=CONV, max_iter=MAXIT)
    peak = mix[0][2]["peak_in_flight"]
    for k, ((Ea, ia, sa), (Em, im, sm)) in enumerate(zip(alone, mix)):
        print(f"request {k}: n={sa['nBasis']} mult={requests[k]['multiplicity']} alone {Ea} ({ia}) "
              f"mix {Em} ({im}) latency {sm['finish'] - sm['arrival']:.3f} s")
        assert sa["converged_steps"] == len(Ea) and sm["converged_steps"] == len(Em)
        assert np.abs(Ea - Em).max() < 1e-9
    print(f"peak in flight {peak}")
    assert peak > 1


def test_memory_budget_serialises(requests):
    out = serve([dict(r, arrival=0.0) for r in requests], slots=4, conv=CONV, mem_budget=1.0, max_iter=MAXIT)
    ref = serve([dict(r, arrival=0.0) for r in requests], slots=4, conv=CONV, max_iter=MAXIT)
    assert out[0][2]["peak_in_flight"] == 1
    for (E1, _, _), (E2, _, _) in zip(out, ref):
        assert np.abs(E1 - E2).max() < 1e-9


def test_steps_match_uhf(requests):
    res = serve(requests, slots=2, conv=CONV, max_iter=MAXIT)
    for r, (E, it, st) in list(zip(requests, res))[:2]:
        Eu = _uhf_steps(r["Z"], r["xyz_bohr"], r["multiplicity"])
        for k in range(len(E)):
            print(f"step {k}: served {E[k]:.10f} ({it[k]} it), uhf() {Eu[k]:.10f}")
            assert Eu[k] != 0.0 and abs(E[k] - Eu[k]) < 1e-8
        assert it[1] <= it[0]


def test_mixed_within_budget(requests):
    e64 = serve(requests, slots=4, conv=CONV, max_iter=MAXIT)
    emx = serve(requests, slots=4, conv=CONV, mixed=1e-4, max_iter=MAXIT)
    worst = max(float(np.abs(a[0] - b[0]).max()) for a, b in zip(e64, emx))
    print(f"mixed vs FP64: max |dE| {worst:.2e} Eh")
    assert worst <= 1e-6
