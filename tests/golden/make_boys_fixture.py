"""Writes tests/golden/boys_mpmath.json: F_m(T) for m = 0..9 by mpmath at 40
significant digits, F_m(T) = T^(-m-1/2) gamma(m+1/2, T) / 2 (lower incomplete
gamma), F_m(0) = 1/(2m+1).  The T grid covers both branches of the device
evaluation: the table range with points on and between grid nodes, points
either side of every per-L switch to the large-T branch, and T up to 80.

Usage: python tests/golden/make_boys_fixture.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))
# per-L switch points of the device evaluation (table rows up to these T)
SWITCH = [36.5, 40.5, 44.0, 47.0, 50.0, 52.5, 55.0, 57.5, 60.0, 62.5]


def boys(m, T):
    if T == 0:
        return mp.mpf(1) / (2 * m + 1)
    T = mp.mpf(T)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, T) / (2 * T ** (m + mp.mpf(1) / 2))


def grid():
    ts = {0.0, 1e-12, 1e-8, 1e-5, 1e-3, 0.01, 0.049, 0.05, 0.051, 0.0999, 0.1}
    k = 0.137
    while k < 80.0:
        ts.add(round(k, 6))
        k *= 1.21
    for t in (1.0, 2.5, 5.0, 7.77, 10.0, 12.34, 20.0, 25.05, 30.0, 33.33):
        ts.add(t)
    for s in SWITCH:
        for d in (-0.06, -0.051, -0.049, -0.03, 0.0, 0.02, 0.1):
            ts.add(round(s + d, 6))
    for t in (59.9, 59.95, 60.0, 60.04, 61.0, 65.0, 70.0, 75.5, 80.0):
        ts.add(t)
    return sorted(ts)


def main():
    T = grid()
    F = [[float(boys(m, t)) for m in range(10)] for t in T]
    out = os.path.join(HERE, "boys_mpmath.json")
    with open(out, "w") as f:
        json.dump({"T": T, "F": F, "mmax": 9, "digits": 40}, f, indent=0)
    print(f"{len(T)} points -> {out}")


if __name__ == "__main__":
    main()
