"""Device Boys function against mpmath (tests/golden/boys_mpmath.json).

sqjk_boys runs the integral kernels' own Boys code (boys<L, R> in
csrc/sq_md.h) on the device for every total angular momentum L <= 9, on a T
grid that straddles each L's switch from the LDS table (Taylor expansion plus
downward recursion) to the large-T branch (sqrt(pi/T)/2 plus upward
recursion without exp(-T)), up to T = 80.  Both the FP64 path and the FP32
path of the mixed-precision kernels are checked.
Tolerances (relative): FP64 2e-14 (the table is built in long double; the
reference itself truncates sqrt(pi)/2 at 4.7e-14); FP32 2e-6.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from sqamd import dptr, lib

pytestmark = pytest.mark.gpu
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "boys_mpmath.json")


def _device_boys(L, T, fp32):
    lib().sqjk_boys.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double)]
    T = np.ascontiguousarray(T, dtype=np.float64)
    F = np.zeros((len(T), L + 1))
    rc = lib().sqjk_boys(L, len(T), dptr(T), int(fp32), 0, dptr(F))
    assert rc == 0, lib().sqjk_errmsg()
    return F


@pytest.mark.parametrize("fp32,tol", [(False, 2e-14), (True, 2e-6)])
@pytest.mark.parametrize("L", list(range(10)))
def test_device_boys_vs_mpmath(L, fp32, tol):
    fx = json.load(open(FIX))
    T = np.array(fx["T"])
    ref = np.array(fx["F"])[:, :L + 1]
    F = _device_boys(L, T, fp32)
    rel = np.abs(F - ref) / np.abs(ref)
    worst = np.unravel_index(np.argmax(rel), rel.shape)
    print(f"L={L} fp32={fp32}: max rel err {rel.max():.2e} at T={T[worst[0]]}, m={worst[1]}")
    assert rel.max() < tol


def test_device_boys_rejects_bad_input():
    T = np.array([-1.0])
    F = np.zeros(2)
    lib().sqjk_boys.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_double)]
    assert lib().sqjk_boys(1, 1, dptr(T), 0, 0, dptr(F)) != 0
    assert lib().sqjk_boys(10, 1, dptr(np.zeros(1)), 0, 0, dptr(F)) != 0
