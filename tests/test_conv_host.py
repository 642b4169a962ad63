"""Host density damping / DIIS drop-ins (src/conv.c:40-844) on the CPU.

The 2- and 3-point closed forms of the reference and its 4-point linear
system are all the constrained least-squares DIIS solution
min |sum_a c_a e_a|^2 subject to sum_a c_a = 1; checked against numpy's
solution of that problem (a KKT system solved independently), and damping
against its definition.  Also: the reset call, the damped first K-1 calls
and the fallback to damping when the newest error grows.
"""
import ctypes as C

import numpy as np
import pytest

from sqamd import dptr, lib

N = 6


def _fn(name):
    f = getattr(lib(), name)
    f.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    f.restype = C.c_double
    return f


def _kkt(errs):
    K = len(errs)
    B = np.array([[np.dot(a.ravel(), b.ravel()) for b in errs] for a in errs])
    A = np.zeros((K + 1, K + 1))
    A[:K, :K] = B
    A[:K, K] = A[K, :K] = -1.0
    rhs = np.zeros(K + 1)
    rhs[K] = -1.0
    return np.linalg.solve(A, rhs)[:K], B


@pytest.mark.parametrize("name,K", [("conv_diis2", 2), ("conv_diis3", 3), ("conv_diis4", 4)])
def test_diis_matches_constrained_least_squares(name, K):
    f = _fn(name)
    f(0, 0.0, None, None)
    rng = np.random.default_rng(K)
    alpha = 0.3
    # a converging sequence: P_t = P* + 0.5^t noise
    Pstar = rng.standard_normal((N, N))
    inputs, errors = [], []
    P = Pstar + rng.standard_normal((N, N))
    PA, PB = P.copy(), P.copy()
    f(N, alpha, dptr(PA), dptr(PB))
    inputs.append(P.copy())
    for t in range(1, K + 3):
        cur = Pstar + 0.5 ** t * rng.standard_normal((N, N))
        PA, PB = cur.copy(), cur.copy()
        rms = f(N, alpha, dptr(PA), dptr(PB))
        e = cur - inputs[-1]
        errors.append(e)
        assert abs(rms - np.sqrt(2 * np.sum(e * e) / N / N / 2)) < 1e-14
        if t <= K - 1:
            expect = alpha * cur + (1 - alpha) * inputs[-1]
        else:
            errs = errors[-K:]
            c, B = _kkt(errs)
            outs = [inputs[-K + a] + errors[-K + a] for a in range(K)]
            expect = sum(c[a] * outs[a] for a in range(K))
            if B[K - 1, K - 1] > B[K - 2, K - 2]:   # the reference's fallback to damping
                expect = alpha * cur + (1 - alpha) * inputs[-1]
        assert np.abs(PA - expect).max() < 1e-10 * max(1.0, np.abs(expect).max()), t
        assert np.array_equal(PA, PB)
        inputs.append(PA.copy())
    f(0, 0.0, None, None)


def test_damping_definition():
    """P <- alpha P + (1 - alpha) P_prev for both spins on every call."""
    f = _fn("conv_damping")
    f(0, 0.0, None, None)
    rng = np.random.default_rng(1)
    prevA = rng.standard_normal((N, N))
    prevB = rng.standard_normal((N, N))
    PA, PB = prevA.copy(), prevB.copy()
    assert f(N, 0.25, dptr(PA), dptr(PB)) == 0.0
    for _ in range(4):
        curA, curB = rng.standard_normal((N, N)), rng.standard_normal((N, N))
        PA, PB = curA.copy(), curB.copy()
        rms = f(N, 0.25, dptr(PA), dptr(PB))
        ref = np.sqrt((np.sum((curA - prevA) ** 2) + np.sum((curB - prevB) ** 2)) / N / N / 2)
        assert abs(rms - ref) < 1e-14
        assert np.abs(PA - (0.25 * curA + 0.75 * prevA)).max() < 1e-14
        assert np.abs(PB - (0.25 * curB + 0.75 * prevB)).max() < 1e-14
        prevA, prevB = PA.copy(), PB.copy()
    f(0, 0.0, None, None)
