"""Molecule and basis input for the engine (host side).

* GAMESS $DATA basis files, expanded per atom in file order with the
  Cartesian component order and primitive-only normalisation of the
  reference (src/basis.c:316-623): norm_k = 1/sqrt(<g_k|g_k>) for each
  Cartesian component, the contraction itself is not renormalised.  The
  self-overlap is used in closed form,
  <g|g> = (pi/2a)^{3/2} (2l-1)!!(2m-1)!!(2n-1)!! / (4a)^{l+m+n}.
* XYZ files in angstrom -> bohr with 1.889725988579 (src/util.h:21).
* Seeded generators for the synthetic benchmark geometries.
"""
import math
import os

import numpy as np

from .abi import Basis, Mol

ANGSTROM2BOHR = 1.889725988579
DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data")
SYMBOLS = ["", "H", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE", "NA", "MG", "AL", "SI",
           "P", "S", "CL", "AR"]
NAMES = ["", "HYDROGEN", "HELIUM", "LITHIUM", "BERYLLIUM", "BORON", "CARBON", "NITROGEN",
         "OXYGEN", "FLUORINE", "NEON", "SODIUM", "MAGNESIUM", "ALUMINUM", "SILICON",
         "PHOSPHORUS", "SULFUR", "CHLORINE", "ARGON"]
COMPONENTS = {
    "S": [(0, 0, 0)], "P": [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    "D": [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
    "F": [(3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (1, 2, 0), (2, 0, 1), (1, 0, 2),
          (0, 2, 1), (0, 1, 2), (1, 1, 1)],
}


def basis_path(name):
    return os.path.join(DATA, "basis", name + ".txt")


def read_gamess_basis(path):
    """-> {Z: [(components, exps, coefs), ...]} in file order."""
    words = open(path).read().split()
    i = words.index("$DATA") + 1
    db, Z = {}, 0

    def num(s):
        return float(s.replace("D", "E").replace("d", "e"))

    while i < len(words) and words[i] != "$END":
        w = words[i].upper()
        if w in NAMES[1:]:
            Z = NAMES.index(w)
            i += 1
            continue
        if w in ("S", "L", "P", "D", "F"):
            nc = int(words[i + 1])
            i += 2
            rows = []
            for _ in range(nc):
                k = 4 if w == "L" else 3
                rows.append([num(x) for x in words[i + 1:i + k]])
                i += k
            rows = np.array(rows)
            if w == "L":
                db.setdefault(Z, []).append((COMPONENTS["S"], rows[:, 0], rows[:, 1]))
                db[Z].append((COMPONENTS["P"], rows[:, 0], rows[:, 2]))
            else:
                db.setdefault(Z, []).append((COMPONENTS[w], rows[:, 0], rows[:, 1]))
            continue
        i += 1
    return db


def _dfact(k):
    r = 1.0
    while k > 1:
        r *= k
        k -= 2
    return r


def prim_norm(a, l, m, n):
    s = (math.pi / (2 * a)) ** 1.5 * _dfact(2 * l - 1) * _dfact(2 * m - 1) * _dfact(2 * n - 1) / (4 * a) ** (l + m + n)
    return 1.0 / math.sqrt(s)


def make_basis(mol, basis):
    """Basis (GTOBasis_t array) for a Mol, basis = name or path."""
    path = basis if os.path.exists(str(basis)) else basis_path(basis)
    db = read_gamess_basis(path)
    funcs = []
    for Z, x, y, z in zip(mol.Z, mol.x, mol.y, mol.z):
        if int(Z) not in db:
            raise ValueError(f"element Z={Z} not in {path}")
        for comps, ex, co in db[int(Z)]:
            for (l, m, n) in comps:
                nr = np.array([prim_norm(a, l, m, n) for a in ex])
                funcs.append((l, m, n, (float(x), float(y), float(z)), ex, co, nr))
    return Basis(funcs)


def read_xyz(path, charge=0):
    lines = open(path).read().splitlines()
    nat = int(lines[0].split()[0])
    Z, xyz = [], []
    for ln in lines[2:2 + nat]:
        w = ln.split()
        Z.append(SYMBOLS.index(w[0].upper()))
        xyz.append([float(v) * ANGSTROM2BOHR for v in w[1:4]])
    return Mol(Z, xyz, charge)


def mol_from_angstrom(Z, xyz_ang, charge=0):
    return Mol(Z, np.asarray(xyz_ang, dtype=np.float64) * ANGSTROM2BOHR, charge)


# ----------------------------------------------------------------- geometries
def water_c1():
    """O at the origin, H at (+-0.757, 0.587, 0) angstrom."""
    return [8, 1, 1], np.array([[0.0, 0.0, 0.0], [0.757, 0.587, 0.0], [-0.757, 0.587, 0.0]])


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster(nmol, seed, min_oo=2.7):
    """Rigid monomers with random orientations, O placed by random sequential
    addition (all O-O >= min_oo angstrom) in a cube sized for 1 g/cm^3."""
    Z0, mono = water_c1()
    side = (nmol * 18.015 / 6.02214076e23 / 1.0) ** (1.0 / 3.0) * 1e8  # angstrom
    for attempt in range(1000):
        rng = np.random.default_rng([seed, attempt])
        centers = []
        for _ in range(nmol):
            for _try in range(20000):
                c = rng.uniform(0, side, 3)
                if all(np.linalg.norm(c - o) >= min_oo for o in centers):
                    centers.append(c)
                    break
            else:
                break
        if len(centers) == nmol:
            break
    else:
        raise RuntimeError("water_cluster: packing failed")
    Z, xyz = [], []
    for c in centers:
        R = _random_rotation(rng)
        for zz, r in zip(Z0, mono):
            Z.append(zz)
            xyz.append(c + R @ r)
    xyz = np.array(xyz)
    return Z, xyz - xyz.mean(axis=0)


def c60(r_single=1.45, r_double=1.40):
    """Truncated icosahedron with pentagon edges r_single and hexagon-fusing
    edges r_double (angstrom), centred at the origin."""
    phi = (1 + 5 ** 0.5) / 2
    # vertices of the truncated icosahedron: even permutations of
    # (0,+-1,+-3phi), (+-1,+-(2+phi),+-2phi), (+-phi,+-2,+-(2phi+1))
    base = [(0, 1, 3 * phi), (1, 2 + phi, 2 * phi), (phi, 2, 2 * phi + 1)]
    pts = set()
    for b in base:
        for perm in [(0, 1, 2), (1, 2, 0), (2, 0, 1)]:
            v = [b[perm[0]], b[perm[1]], b[perm[2]]]
            for sx in (1, -1):
                for sy in (1, -1):
                    for sz in (1, -1):
                        pts.add((round(sx * v[0], 9), round(sy * v[1], 9), round(sz * v[2], 9)))
    P = np.array(sorted(pts))
    assert len(P) == 60
    # unit edge length 2; classify edges: pentagon edges lie on 5-rings
    d = np.linalg.norm(P[:, None] - P[None], axis=-1)
    edges = [(i, j) for i in range(60) for j in range(i + 1, 60) if abs(d[i, j] - 2.0) < 1e-6]
    nb = {i: [j for j in range(60) if abs(d[i, j] - 2.0) < 1e-6] for i in range(60)}

    def in_pentagon(i, j):
        # edge (i,j) is on a pentagon iff a 5-cycle passes through it
        for k in nb[j]:
            if k == i:
                continue
            for m in nb[k]:
                if m in (i, j):
                    continue
                for q in nb[m]:
                    if q in (i, j, k):
                        continue
                    if i in nb[q]:
                        return True
        return False
    # Every atom lies on exactly one pentagon.  Shrink each pentagon about its
    # centroid by t (bisection on the bond-length ratio), then scale the cage
    # so the pentagon edges have length r_single.
    pent = [(i, j) for (i, j) in edges if in_pentagon(i, j)]
    pent_set = set(pent)
    Pn = P / 2.0
    ring = {i: [j for j in nb[i] if (min(i, j), max(i, j)) in pent_set] for i in range(60)}
    # pentagon centroid of atom i: average of the 5-ring through i
    def pent_members(i):
        mem = {i}
        frontier = [i]
        while frontier:
            a = frontier.pop()
            for b in ring[a]:
                if b not in mem:
                    mem.add(b)
                    frontier.append(b)
        return sorted(mem)
    cent = np.array([Pn[pent_members(i)].mean(axis=0) for i in range(60)])

    def geom(t, s):
        return s * (cent + t * (Pn - cent))

    def lengths(X):
        ls = np.mean([np.linalg.norm(X[i] - X[j]) for (i, j) in pent])
        ld = np.mean([np.linalg.norm(X[i] - X[j]) for (i, j) in edges if (i, j) not in set(pent)])
        return ls, ld
    lo, hi = 0.5, 1.5
    for _ in range(200):
        t = 0.5 * (lo + hi)
        ls, ld = lengths(geom(t, 1.0))
        if ls / ld < r_single / r_double:
            lo = t
        else:
            hi = t
    ls, ld = lengths(geom(t, 1.0))
    X = geom(t, r_single / ls)
    return [6] * 60, X - X.mean(axis=0)


# ------------------------------------------------ small-molecule instances
def _place(center, direction, r):
    d = np.asarray(direction, dtype=float)
    return np.asarray(center, dtype=float) + r * d / np.linalg.norm(d)


def small_molecule(name):
    """Equilibrium-like geometries (angstrom) of the small molecules of the
    batched workload, built from standard bond lengths and angles:
    H2O, NH3, CH4, C2H2, C2H4, C2H6, CH2O, CH3OH, C6H6."""
    t = np.deg2rad
    if name == "H2O":
        return water_c1()
    if name == "NH3":  # N-H 1.012, H-N-H 106.7
        r, a = 1.012, t(106.7)
        # angle from the C3 axis for a given H-N-H angle
        th = np.arcsin(np.sqrt((1 - np.cos(a)) / 1.5))
        H = [[r * np.sin(th) * np.cos(k * 2 * np.pi / 3), r * np.sin(th) * np.sin(k * 2 * np.pi / 3),
              -r * np.cos(th)] for k in range(3)]
        return [7, 1, 1, 1], np.array([[0, 0, 0]] + H)
    if name == "CH4":  # C-H 1.087, tetrahedral
        d = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3) * 1.087
        return [6, 1, 1, 1, 1], np.vstack([[0, 0, 0], d])
    if name == "C2H2":  # C-C 1.203, C-H 1.063
        return [6, 6, 1, 1], np.array([[0, 0, 0.6015], [0, 0, -0.6015], [0, 0, 1.6645], [0, 0, -1.6645]])
    if name == "C2H4":  # C=C 1.339, C-H 1.086, H-C-H 117.6
        h = t(117.6) / 2
        z, hx, hz = 0.6695, 1.086 * np.sin(h), 1.086 * np.cos(h)
        return [6, 6, 1, 1, 1, 1], np.array([[0, 0, z], [0, 0, -z], [hx, 0, z + hz], [-hx, 0, z + hz],
                                             [hx, 0, -z - hz], [-hx, 0, -z - hz]])
    if name == "C2H6":  # C-C 1.535, C-H 1.094, H-C-C 111.2, staggered
        z, r, a = 0.7675, 1.094, t(180 - 111.2)
        H = []
        for s, ph0 in ((1, 0.0), (-1, np.pi / 3)):
            for k in range(3):
                ph = ph0 + k * 2 * np.pi / 3
                H.append([r * np.sin(a) * np.cos(ph), r * np.sin(a) * np.sin(ph), s * (z + r * np.cos(a))])
        return [6, 6] + [1] * 6, np.array([[0, 0, z], [0, 0, -z]] + H)
    if name == "CH2O":  # C=O 1.205, C-H 1.111, H-C-H 116.1
        h = t(116.1) / 2
        return [6, 8, 1, 1], np.array([[0, 0, 0], [0, 0, 1.205], [1.111 * np.sin(h), 0, -1.111 * np.cos(h)],
                                       [-1.111 * np.sin(h), 0, -1.111 * np.cos(h)]])
    if name == "CH3OH":  # C-O 1.427, O-H 0.956, C-H 1.094, C-O-H 108.9
        a = t(108.9)
        O = np.array([0, 0, 1.427])
        Ho = O + 0.956 * np.array([np.sin(np.pi - a), 0, np.cos(np.pi - a)])
        Z, X = small_molecule("CH4")
        Hs = [X[2], X[3], X[4]]  # three of the tetrahedral directions, rotated to point away from O
        R = np.eye(3)
        v = X[1] / np.linalg.norm(X[1])
        # rotate the tetrahedron so that the first C-H direction points along +z (taken by O)
        w = np.cross(v, [0, 0, 1])
        s, c = np.linalg.norm(w), np.dot(v, [0, 0, 1])
        if s > 1e-12:
            k = w / s
            Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
            R = np.eye(3) + s * Kx + (1 - c) * Kx @ Kx
        Hs = [R @ h for h in Hs]
        return [6, 8, 1, 1, 1, 1], np.vstack([[0, 0, 0], O, Ho, Hs])
    if name == "C6H6":  # C-C 1.397, C-H 1.084
        C = [[1.397 * np.cos(k * np.pi / 3), 1.397 * np.sin(k * np.pi / 3), 0] for k in range(6)]
        H = [[2.481 * np.cos(k * np.pi / 3), 2.481 * np.sin(k * np.pi / 3), 0] for k in range(6)]
        return [6] * 6 + [1] * 6, np.array(C + H)
    raise ValueError(name)


BATCH_MOLECULES = ["H2O", "NH3", "CH4", "C2H2", "C2H4", "C2H6", "CH2O", "CH3OH", "C6H6"]
BATCH_BASES = ["sto3g", "631g", "631g_star"]


def batch_instances(count=256, seed=4256, uhf_fraction=0.25, sigma=0.05):
    """Seeded instances of the batched workload: molecule and basis drawn
    uniformly from BATCH_MOLECULES x BATCH_BASES (7 to 102 basis
    functions), every coordinate displaced by N(0, sigma) angstrom, and a
    fraction turned into UHF doublet cations (charge +1, multiplicity 2).
    Returns dicts with name, basis, mol (Mol, bohr), charge, nEA, nEB,
    restricted."""
    rng = np.random.default_rng(seed)
    nuhf = int(round(uhf_fraction * count))
    uhf = set(rng.choice(count, nuhf, replace=False).tolist())
    out = []
    for i in range(count):
        name = BATCH_MOLECULES[rng.integers(len(BATCH_MOLECULES))]
        basis = BATCH_BASES[rng.integers(len(BATCH_BASES))]
        Z, X = small_molecule(name)
        X = np.asarray(X, dtype=float) + rng.normal(0.0, sigma, np.shape(X))
        charge = 1 if i in uhf else 0
        ne = int(np.sum(Z)) - charge
        mult = 2 if i in uhf else 1
        out.append({"name": name, "basis": basis, "mol": mol_from_angstrom(Z, X, charge), "charge": charge,
                    "nEA": (ne + mult - 1) // 2, "nEB": (ne - mult + 1) // 2, "restricted": i not in uhf})
    return out


# ------------------------------------------------------- served requests
def random_chain(natom, rng, bond=(1.0, 1.5), h_bond=(1.0, 1.1), min_heavy=2.0, min_any=1.5):
    """A self-avoiding chain of heavy atoms (C 70 %, N 15 %, O 15 %) with
    bond lengths U(bond) angstrom, hydrogens on the chain atoms at U(h_bond)
    (at most three per atom); non-bonded heavy atoms at least `min_heavy`
    apart and every other non-bonded pair at least `min_any`.  About 60 % of
    the atoms are heavy.  Returns (Z, xyz angstrom)."""
    nheavy = max(1, int(round(0.6 * natom)))
    nh = natom - nheavy
    for _attempt in range(200):
        Z = [int(rng.choice([6, 7, 8], p=[0.7, 0.15, 0.15])) for _ in range(nheavy)]
        X = [np.zeros(3)]
        prev_dir = None
        ok = True
        for k in range(1, nheavy):
            for _try in range(200):
                d = rng.standard_normal(3)
                d /= np.linalg.norm(d)
                if prev_dir is not None and np.dot(d, prev_dir) < 0.0:  # keep the chain from folding back
                    continue
                p = X[-1] + d * rng.uniform(*bond)
                if all(np.linalg.norm(p - q) >= min_heavy for q in X[:-1]):
                    X.append(p)
                    prev_dir = d
                    break
            else:
                ok = False
                break
        if not ok:
            continue
        count = [0] * nheavy
        H = []
        for _ in range(nh):
            for _try in range(500):
                a = int(rng.integers(nheavy))
                if count[a] >= 3:
                    continue
                d = rng.standard_normal(3)
                d /= np.linalg.norm(d)
                p = X[a] + d * rng.uniform(*h_bond)
                if all(np.linalg.norm(p - q) >= min_any for i, q in enumerate(X) if i != a) and \
                        all(np.linalg.norm(p - q) >= min_any for q in H):
                    H.append(p)
                    count[a] += 1
                    break
            else:
                ok = False
                break
        if ok:
            return Z + [1] * nh, np.array(X + H)
    raise RuntimeError("random_chain: placement failed")


def serving_requests(count=16, rate=4.0, seed=5005, steps=2, atoms=(10, 200), sigma=0.05):
    """Seeded requests of the serving workload: atom counts log-uniform in
    `atoms`, random C/H/N/O chains (random_chain), `steps` geometry steps
    each displaced by N(0, sigma) angstrom from the base geometry, Poisson
    arrivals at `rate` per second.  RHF for an even electron count, else a
    UHF doublet.  Returns dicts with Z, xyz_bohr [steps, natom, 3], charge,
    multiplicity, arrival (s), natom."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += rng.exponential(1.0 / rate) if i > 0 else 0.0
        natom = int(round(np.exp(rng.uniform(np.log(atoms[0]), np.log(atoms[1])))))
        Z, X = random_chain(natom, rng)
        geo = np.stack([(X + rng.normal(0.0, sigma, X.shape)) * ANGSTROM2BOHR for _ in range(steps)])
        ne = int(np.sum(Z))
        out.append({"Z": np.array(Z, dtype=np.int32), "xyz_bohr": geo, "charge": 0,
                    "multiplicity": 1 if ne % 2 == 0 else 2, "arrival": t, "natom": natom})
    return out
