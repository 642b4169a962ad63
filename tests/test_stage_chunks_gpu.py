"""Exchange staging held in ket-pair chunks.

A context whose pair-indexed exchange staging would exceed its budget
(SQJK_STAGE_MB at creation, 256 MiB by default) runs the launches of each
ket pair class chunk by chunk through one buffer per spin, folding every
chunk into the half-accumulators before the next.  The result is the same
J/K as whole-class staging (same integrals, summed in another order):
1e-11 max-abs, RHF and UHF; (H2O)4 against the oracle (the CPU restatement
of src/matrix.c:3773-3986) at 1e-10.  The device-memory estimate of a
chunked context is below that of a whole-class one.
"""
import ctypes as C
import os

import numpy as np
import pytest

from molecules import cluster, random_density
from sqamd import JKEngine, lib

pytestmark = pytest.mark.gpu


def _with_budget(mb, fn):
    old = os.environ.get("SQJK_STAGE_MB")
    os.environ["SQJK_STAGE_MB"] = str(mb)
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["SQJK_STAGE_MB"]
        else:
            os.environ["SQJK_STAGE_MB"] = old


def _footprint(bas, uhf):
    L = lib()
    L.sqjk_footprint_bytes.restype = C.c_longlong
    return L.sqjk_footprint_bytes(bas.n, bas.ptr, int(uhf), 0)


@pytest.mark.parametrize("restricted", [True, False])
def test_chunked_equals_whole_class(restricted):
    mol, bas = cluster(16, 1601)
    n = bas.n
    PA = random_density(n, 31, 0.01)
    PB = PA if restricted else random_density(n, 32, 0.01)
    whole = _with_budget(100000, lambda: JKEngine(bas.ptr, n))
    chunked = _with_budget(8, lambda: JKEngine(bas.ptr, n))
    r0 = whole.build(PA, PB, 1e-11, 1e-11, restricted)
    r1 = chunked.build(PA, PB, 1e-11, 1e-11, restricted)
    k0, k1 = whole.stats()["kernel_ms"], chunked.stats()["kernel_ms"]
    whole.close()
    chunked.close()
    d = max(float(np.abs(a - b).max()) for a, b in zip(r0, r1))
    f_whole = _with_budget(100000, lambda: _footprint(bas, not restricted))
    f_chunk = _with_budget(8, lambda: _footprint(bas, not restricted))
    print(f"(H2O)16 n={n} restricted={restricted}: max diff {d:.2e}; kernel {k0:.2f} ms whole, {k1:.2f} ms "
          f"chunked; footprint {f_whole / 2**20:.0f} -> {f_chunk / 2**20:.0f} MiB")
    assert d < 1e-11
    assert f_chunk < f_whole


def test_chunked_against_oracle(oracle):
    mol, bas = cluster(4, 401)
    n = bas.n
    PA, PB = random_density(n, 41, 0.05), random_density(n, 42, 0.05)
    eng = _with_budget(1, lambda: JKEngine(bas.ptr, n))
    J, KA, KB = eng.build(PA, PB, 1e-10, 1e-10, False)
    eng.close()
    sch = oracle.schwarz(bas.ptr, n)
    Jo, KAo, KBo = oracle.jk(oracle.jk_context(False, 0), bas.ptr, n, PA, PB, sch, 1e-10, 1e-10, False)
    d = max(float(np.abs(J - Jo).max()), float(np.abs(KA - KAo).max()), float(np.abs(KB - KBo).max()))
    print(f"(H2O)4 UHF, 1 MiB staging budget: vs oracle {d:.2e}")
    assert d < 1e-10
