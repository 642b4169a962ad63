"""Algorithmic FP64 work per primitive quartet, per angular class.

Counted from the loop structure of the two kernels (FMA = 2 flops,
multiply/add = 1, division and square root = 1, exp = 1):
  per-thread kernel (sq_jk_dev.h) for classes with <= 81 integrals per
  quartet: geometry, Boys (9-term Taylor + downward recursion), Hermite
  R tensor (level recursion), ket Hermite coefficients, the W transform
  and the bra contraction, summed over the bra-component slices;
  cooperative kernel (sq_jk_wave.h) above 81: Boys per order, closed-form
  R entries, Hermite tables, W and the contraction.
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


def _slices(la, lb, lc, ld):
    nab, ncd = ncart(la) * ncart(lb), ncart(lc) * ncart(ld)
    sz = max(1, min(36 // ncd, nab)) if ncd <= 36 else 1
    return [list(range(s, min(s + sz, nab))) for s in range(0, nab, sz)]


@lru_cache(maxsize=None)
def flops_per_prim_quartet(la, lb, lc, ld):
    L = la + lb + lc + ld
    lab = la + lb
    nb, nc, nd = ncart(lb), ncart(lc), ncart(ld)
    nint = ncart(la) * nb * nc * nd
    if nint <= 81:
        total = 0
        for sl in _slices(la, lb, lc, ld):
            f = 20 + (28 + 3 * L) + _R_level_flops(L) + 3 * _herm_coef_flops(lc, ld)
            comps = [(cart(la, ab // nb), cart(lb, ab % nb)) for ab in sl]
            mx = max(a[0] + b[0] for a, b in comps)
            my = max(a[1] + b[1] for a, b in comps)
            mz = max(a[2] + b[2] for a, b in comps)
            for cd in range(nc * nd):
                c, d = cart(lc, cd // nd), cart(ld, cd % nd)
                kx, ky, kz = c[0] + d[0], c[1] + d[1], c[2] + d[2]
                nw = sum(1 for t in range(mx + 1) for u in range(my + 1) for v in range(mz + 1) if t + u + v <= lab)
                f += nw * ((kx + 1) * (ky + 1) * (2 * (kz + 1)) + 3 * (kx + 1) * (ky + 1))
                for a, b in comps:
                    nx, ny, nz = a[0] + b[0], a[1] + b[1], a[2] + b[2]
                    f += (nx + 1) * (ny + 1) * (2 * (nz + 1)) + 3 * (nx + 1) * (ny + 1) + 2
            total += f
        return total
    # cooperative kernel
    f = 20 + (L + 1) * 30
    f += 3 * _herm_coef_flops(la, lb) + 3 * _herm_coef_flops(lc, ld)
    for T in range(L + 1):
        for U in range(L + 1 - T):
            for V in range(L + 1 - T - U):
                terms = (T // 2 + 1) * (U // 2 + 1) * (V // 2 + 1)
                f += terms * 2 + (T + U + V)
    for cd in range(nc * nd):
        c, d = cart(lc, cd // nd), cart(ld, cd % nd)
        kx, ky, kz = c[0] + d[0], c[1] + d[1], c[2] + d[2]
        f += nherm(lab) * ((kx + 1) * (ky + 1) * (2 * (kz + 1)) + 3 * (kx + 1) * (ky + 1))
    for ab in range(ncart(la) * nb):
        a, b = cart(la, ab // nb), cart(lb, ab % nb)
        nx, ny, nz = a[0] + b[0], a[1] + b[1], a[2] + b[2]
        f += nc * nd * ((nx + 1) * (ny + 1) * (2 * (nz + 1)) + 3 * (nx + 1) * (ny + 1) + 2)
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
