"""Algorithmic work per primitive quartet, per angular class.

The count is the one-pass McMurchie-Davidson evaluation that the kernels
implement (FMA = 2 flops, multiply/add = 1, division and square root = 1,
exp = 1), independent of how a kernel splits a quartet over lanes:
geometry and prefactor, Boys (9-term Taylor from the table + downward
recursion), the Hermite Coulomb tensor R (level recursion), the ket Hermite
coefficients, W[cd][tuv] = sum E^cd R for every ket component, and the bra
contraction for every integral.  Work that a kernel repeats (R and Boys per
ket-component slice, bra Hermite coefficients per bra primitive) is not
counted, so achieved/peak is a fraction of useful work.
Digestion is counted per contracted quartet: 12 flops per integral (RHF),
20 (UHF).  The table is a model; the rocprofv3 FP64 instruction counters of
the same run are its cross-check.
"""
from functools import lru_cache


def ncart(l):
    return (l + 1) * (l + 2) // 2


def cart(l, i):
    if l == 0:
        return (0, 0, 0)
    if l == 1:
        return [(1, 0, 0), (0, 1, 0), (0, 0, 1)][i]
    return [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)][i]


def nherm(L):
    return (L + 1) * (L + 2) * (L + 3) // 6


def _herm_coef_flops(li, lj):
    f = 0
    for i in range(li):
        f += sum(2 * ((t >= 1) + (t <= i) + (t + 1 <= i)) for t in range(i + 2))
    for i in range(li + 1):
        for j in range(lj):
            f += sum(2 * ((t >= 1) + (t <= i + j) + (t + 1 <= i + j)) for t in range(i + j + 2))
    return f


def _R_level_flops(L):
    f = 2 * (L + 1)
    for n in range(L - 1, -1, -1):
        for o in range(1, L - n + 1):
            for m in range(o + 1):
                for v in range(m + 1):
                    u, t = m - v, o - m
                    k = t if t > 0 else (u if u > 0 else v)
                    f += 1 + (2 if k > 1 else 0)
    return f


@lru_cache(maxsize=None)
def flops_per_prim_quartet(la, lb, lc, ld):
    L = la + lb + lc + ld
    lab = la + lb
    nab, nb, nc, nd = ncart(la) * ncart(lb), ncart(lb), ncart(lc), ncart(ld)
    f = 20 + (28 + 3 * L) + _R_level_flops(L) + 3 * _herm_coef_flops(lc, ld)
    for cd in range(nc * nd):
        c, d = cart(lc, cd // nd), cart(ld, cd % nd)
        kx, ky, kz = c[0] + d[0], c[1] + d[1], c[2] + d[2]
        f += nherm(lab) * ((kx + 1) * (ky + 1) * (2 * (kz + 1)) + 3 * (kx + 1) * (ky + 1))
        for ab in range(nab):
            a, b = cart(la, ab // nb), cart(lb, ab % nb)
            nx, ny, nz = a[0] + b[0], a[1] + b[1], a[2] + b[2]
            f += (nx + 1) * (ny + 1) * (2 * (nz + 1)) + 3 * (nx + 1) * (ny + 1) + 2
    return f


def digestion_flops(la, lb, lc, ld, uhf=False):
    return (20 if uhf else 12) * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)


def build_flops(class_stats, uhf=False):
    """Total modelled flops of one build from JKEngine.class_stats()."""
    tot = 0
    for lq, nq, npq in class_stats:
        tot += npq * flops_per_prim_quartet(*lq) + nq * digestion_flops(*lq, uhf=uhf)
    return tot


def integrals(class_stats):
    return sum(nq * ncart(a) * ncart(b) * ncart(c) * ncart(d) for (a, b, c, d), nq, _ in class_stats)


def nint(la, lb, lc, ld):
    """Contracted Cartesian integrals per quartet of the class."""
    return ncart(la) * ncart(lb) * ncart(lc) * ncart(ld)


def class_block(la, lb, lc, ld):
    """Workgroup size of the per-thread class kernel (class_block in
    csrc/sq_jk_dev.h): 512 lanes when a lane holds at most 9 integrals of its
    ket-component slice, else 256."""
    nab, ncd = ncart(la) * ncart(lb), ncart(lc) * ncart(ld)
    budget = 18 if (la, lb, lc, ld) in ((2, 1, 1, 0), (1, 1, 1, 1)) else 36  # slice_budget in csrc/sq_jk_dev.h
    per = max(budget // nab, 1)
    nslice = (ncd + per - 1) // per
    sz = (ncd + nslice - 1) // nslice
    return 512 if nab * sz <= 9 else 256
