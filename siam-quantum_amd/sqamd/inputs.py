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
VALENCE = {6: 4, 7: 3, 8: 2}


def _h_directions(bonds, nh, rng):
    """Unit vectors of nh hydrogens on an atom whose other bonds point along
    `bonds` (unit vectors): near-tetrahedral positions in a frame fixed by
    the bonds (a random azimuth when the frame is not fixed)."""
    if nh == 0:
        return []
    tet = np.deg2rad(109.47)
    def perp(u):
        v = rng.standard_normal(3)
        v -= np.dot(v, u) * u
        return v / np.linalg.norm(v)
    if not bonds:  # lone atom: a tetrahedron
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        return ([u] + _h_directions([u], nh - 1, rng))[:nh]
    if len(bonds) == 1:  # three positions at the tetrahedral angle around the bond
        u = bonds[0]
        e1 = perp(u)
        e2 = np.cross(u, e1)
        phi = rng.uniform(0, 2 * np.pi)
        out = [np.cos(tet) * u + np.sin(tet) * (np.cos(phi + 2 * np.pi * k / 3) * e1 +
                                                 np.sin(phi + 2 * np.pi * k / 3) * e2) for k in range(3)]
        return out[:nh]
    # two bonds: along the outer bisector (one H) or above and below the bond plane (two H)
    u1, u2 = bonds
    b = -(u1 + u2)
    b = b / np.linalg.norm(b) if np.linalg.norm(b) > 1e-6 else perp(u1)
    n = np.cross(u1, u2)
    n = n / np.linalg.norm(n) if np.linalg.norm(n) > 1e-6 else perp(b)
    if nh == 1:
        return [b]
    half = tet / 2
    return [np.cos(half) * b + np.sin(half) * n, np.cos(half) * b - np.sin(half) * n][:nh]


BOND = {(6, 6): 1.53, (6, 7): 1.47, (6, 8): 1.43}
H_BOND = {6: 1.09, 7: 1.01, 8: 0.96}


def random_chain(natom, rng, min_heavy=2.3, min_any=1.6, radical=False):
    """A saturated open chain of about `natom` atoms, built like an alkane
    with ether and amine links: heavy atoms C (70 %), N (15 %), O (15 %), no
    two heteroatoms bonded; bond lengths C-C 1.53, C-N 1.47, C-O 1.43
    angstrom (+-0.03); bond angles 109.5 +- 4 degrees; dihedrals anti (60 %)
    or gauche (+-60), +-10 degrees; every heavy atom completed with
    hydrogens to its valence (C 4, N 3, O 2) on near-tetrahedral positions
    (C-H 1.09, N-H 1.01, O-H 0.96 +- 0.02).  Non-bonded heavy atoms stay
    `min_heavy` apart, every other non-bonded pair `min_any`.  With
    `radical`, one hydrogen is left out (an odd electron count).  Returns
    (Z, xyz angstrom)."""
    # a saturated chain carries about 1.55 hydrogens per heavy atom
    nheavy = max(2, int(round(natom / 2.55)))
    theta = np.deg2rad(180.0 - 109.5)

    def unit(v):
        return v / np.linalg.norm(v)

    def hydrogens(a, X, H, nbr, Z):
        # the hydrogens of heavy atom a, or None if one comes too close to another atom
        out = []
        for d in _h_directions([unit(q - X[a]) for q in nbr], max(0, VALENCE[Z[a]] - len(nbr)), rng):
            p = X[a] + d * (H_BOND[Z[a]] + rng.uniform(-0.02, 0.02))
            if not (all(np.linalg.norm(p - q) >= min_any for i, q in enumerate(X) if i != a) and
                    all(np.linalg.norm(p - q) >= min_any for q in H + out)):
                return None
            out.append(p)
        return out

    def bond(z1, z2):
        return BOND[(min(z1, z2), max(z1, z2))] + rng.uniform(-0.03, 0.03)

    for _attempt in range(50):
        Z = []
        for k in range(nheavy):
            z = int(rng.choice([6, 7, 8], p=[0.7, 0.15, 0.15]))
            Z.append(6 if (k > 0 and Z[-1] != 6 and z != 6) else z)
        X, H = [np.zeros(3)], []
        ok = True
        # grow the chain; atom k-1's hydrogens are placed once atom k fixes its second bond
        for k in range(1, nheavy):
            for _try in range(100):
                if k == 1:
                    d = unit(rng.standard_normal(3))
                else:
                    u = unit(X[k - 1] - X[k - 2])
                    ref = X[k - 3] - X[k - 2] if k >= 3 else rng.standard_normal(3)
                    e1 = ref - np.dot(ref, u) * u
                    e1 = unit(e1) if np.linalg.norm(e1) > 1e-6 else unit(np.cross(u, rng.standard_normal(3)))
                    e2 = np.cross(u, e1)
                    phi = np.deg2rad(rng.choice([180.0, 60.0, -60.0], p=[0.6, 0.2, 0.2]) + rng.uniform(-10, 10))
                    th = theta + np.deg2rad(rng.uniform(-4, 4))
                    d = np.cos(th) * u + np.sin(th) * (np.cos(phi) * e1 + np.sin(phi) * e2)
                p = X[-1] + d * bond(Z[k - 1], Z[k])
                if not (all(np.linalg.norm(p - q) >= min_heavy for q in X[:-1]) and
                        all(np.linalg.norm(p - q) >= min_any for q in H)):
                    continue
                nbr = ([X[k - 2]] if k >= 2 else []) + [p]
                hs = hydrogens(k - 1, X + [p], H, nbr, Z)
                if hs is not None:
                    X.append(p)
                    H += hs
                    break
            else:
                ok = False
                break
        if not ok:
            continue
        for _try in range(50):  # the last atom's hydrogens: a random azimuth about its bond
            hs = hydrogens(nheavy - 1, X, H, [X[-2]], Z)
            if hs is not None:
                H += hs
                break
        else:
            continue
        if radical and H:
            H.pop(int(rng.integers(len(H))))
        return Z + [1] * len(H), np.array(X + H)
    raise RuntimeError("random_chain: placement failed")


def serving_requests(count=16, rate=4.0, seed=5005, steps=2, atoms=(10, 200), sigma=0.05, radicals=0.25):
    """Seeded requests of the serving workload: atom counts log-uniform in
    `atoms`, saturated C/H/N/O chains (random_chain), a fraction
    `radicals` of them with one hydrogen left out, `steps` geometry steps
    each displaced by N(0, sigma) angstrom from the base geometry, Poisson
    arrivals at `rate` per second.  RHF for an even electron count, else a
    UHF doublet.  Returns dicts with Z, xyz_bohr [steps, natom, 3], charge,
    multiplicity, arrival (s), natom."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += rng.exponential(1.0 / rate) if i > 0 else 0.0
        natom = int(round(np.exp(rng.uniform(np.log(atoms[0]), np.log(atoms[1])))))
        Z, X = random_chain(natom, rng, radical=bool(rng.uniform() < radicals))
        geo = np.stack([(X + rng.normal(0.0, sigma, X.shape)) * ANGSTROM2BOHR for _ in range(steps)])
        ne = int(np.sum(Z))
        out.append({"Z": np.array(Z, dtype=np.int32), "xyz_bohr": geo, "charge": 0,
                    "multiplicity": 1 if ne % 2 == 0 else 2, "arrival": t, "natom": len(Z)})
    return out
