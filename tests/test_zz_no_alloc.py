"""The allocation-free J/K build (sqjk_build_device_noalloc) against the
device build it is a variant of (sqjk_build_device), through the C-ABI.

Two calls in a row share the library's scratch: (H2O)4 6-31G* UHF
(76 functions) and then water 6-31G* RHF (19 functions); each must equal
the sqjk_build_device result on the same context to 1e-12 (same kernels,
only the order of the digestion atomics differs), and the water build also
the oracle (CPU restatement of src/matrix.c:3773-3986) to 1e-10.  A build
too large for the scratch returns SQJK_ENOMEM and names its size.
Device buffers come from the HIP runtime (hipMalloc/hipMemcpy) directly.
"""
import ctypes as C
import os

import numpy as np
import pytest

from molecules import cluster, g2, random_density
from sqamd import JKEngine, lib

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


def _hip():
    for name in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    pytest.skip("HIP runtime library not found")


class Dev:
    """n x n double matrices in device memory."""

    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(n * n * 8)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0

    def get(self):
        a = np.zeros((self.n, self.n))
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(a.nbytes), D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _bind():
    L = lib()
    PD = C.c_void_p
    sig = [C.c_void_p, PD, PD, C.c_double, C.c_double, C.c_int, PD, PD, PD]
    L.sqjk_build_device.argtypes = sig
    L.sqjk_build_device_noalloc.argtypes = sig
    L.sqjk_sync.argtypes = [C.c_void_p]
    L.sqjk_errmsg.restype = C.c_char_p
    return L


def _run(L, hip, eng, fn, PA, PB, cut, restricted):
    n = eng.n
    bufs = [Dev(hip, n) for _ in range(5)]
    bufs[0].put(PA)
    bufs[1].put(PB)
    rc = getattr(L, fn)(eng._h, bufs[0].p, bufs[1].p, cut, cut, int(restricted), bufs[2].p, bufs[3].p, bufs[4].p)
    msg = L.sqjk_errmsg().decode() if rc else ""
    if rc == 0:
        assert L.sqjk_sync(eng._h) == 0
        out = tuple(b.get() for b in bufs[2:])
    else:
        out = None
    for b in bufs:
        b.free()
    return rc, msg, out


def test_noalloc_matches_device_build(oracle):
    L, hip = _bind(), _hip()
    cases = [("(H2O)4 6-31G* UHF", cluster(4, 401), False), ("water 6-31G* RHF", g2("h2o_q0m1", "631g_star"), True)]
    for k, (label, (mol, bas), restricted) in enumerate(cases):
        n = bas.n
        PA = random_density(n, 11 + k, 0.05)
        PB = PA if restricted else random_density(n, 21 + k, 0.05)
        eng = JKEngine(bas.ptr, n)
        print(f"call {k}: {label}, n={n}", flush=True)
        rc0, m0, ref = _run(L, hip, eng, "sqjk_build_device", PA, PB, 1e-10, restricted)
        rc1, m1, got = _run(L, hip, eng, "sqjk_build_device_noalloc", PA, PB, 1e-10, restricted)
        assert rc0 == 0, m0
        assert rc1 == 0, m1
        d = max(float(np.abs(a - b).max()) for a, b in zip(got, ref))
        print(f"  noalloc vs device build: max diff {d:.2e}", flush=True)
        assert d < 1e-12
        if n <= 40:
            sch = oracle.schwarz(bas.ptr, n)
            Jo, KAo, KBo = oracle.jk(oracle.jk_context(False, 0), bas.ptr, n, PA, PB, sch, 1e-10, 1e-10, restricted)
            do = max(float(np.abs(got[0] - Jo).max()), float(np.abs(got[1] - KAo).max()),
                     float(np.abs(got[2] - KBo).max()))
            print(f"  noalloc vs oracle: max diff {do:.2e}", flush=True)
            assert do < 1e-10
        eng.close()


def test_noalloc_rejects_builds_above_scratch():
    L, hip = _bind(), _hip()
    mol, bas = cluster(16, 1601)
    n = bas.n
    P = random_density(n, 5, 0.01)
    eng = JKEngine(bas.ptr, n)
    rc, msg, _ = _run(L, hip, eng, "sqjk_build_device_noalloc", P, P, 1e-10, True)
    eng.close()
    print(f"(H2O)16, n={n}: rc {rc}: {msg}")
    assert rc == -5 and "MiB of scratch" in msg
