"""Write the data fixtures this repository carries, from the reference tree.

Run once in a container that has the reference mounted (it is never read at
test or bench time):

  python tests/golden/make_fixtures.py [/root/reference]

Outputs (all data, no code):
  siam-quantum_amd/data/basis/{sto3g,321g,631g,631g_star}.txt
      GAMESS $DATA records for H..Cl copied from basis/*.txt (public Basis Set
      Exchange data; the format is the one read by src/basis.c:316-514).
  tests/golden/g2/<mol>.xyz
      the G2 geometries of examples/G2/*.xyz.
  tests/golden/g2_energies.json
      final SCF energies, Schwarz cut-offs and iteration counts printed in
      docs/validation/sq/<mol>_uhf_<q><m>_{321g,631g_star}.log
      ("Done SCF Loop Total Energy is ..." and
       "Schwarz inequality screening cut-off ..."), plus the first
      "FINAL UHF ENERGY" of docs/validation/gamess/<mol>_uhf_631g_star.log
      (GAMESS UHF/6-31G* at the same input geometry).
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
ELEMENTS = ["HYDROGEN", "HELIUM", "LITHIUM", "BERYLLIUM", "BORON", "CARBON",
            "NITROGEN", "OXYGEN", "FLUORINE", "NEON", "SODIUM", "MAGNESIUM",
            "ALUMINUM", "SILICON", "PHOSPHORUS", "SULFUR", "CHLORINE"]


def cut_basis(src, dst):
    lines = open(src).read().splitlines()
    out, keep, inside = [], False, False
    for ln in lines:
        w = ln.split()
        if not inside:
            if w and w[0] == "$DATA":
                inside = True
                out.append("$DATA")
            continue
        if w and w[0] == "$END":
            break
        if len(w) == 1 and w[0].isalpha() and w[0] not in ("S", "L", "P", "D", "F"):
            keep = w[0].upper() in ELEMENTS
            if keep:
                out.append("  " + w[0].upper())
            continue
        if keep and w:
            out.append(ln.rstrip())
    out.append("$END")
    open(dst, "w").write("\n".join(out) + "\n")


def main():
    for name in ("sto3g", "321g", "631g", "631g_star"):
        cut_basis(os.path.join(REF, "basis", name + ".txt"),
                  os.path.join(REPO, "siam-quantum_amd", "data", "basis", name + ".txt"))
    g2 = os.path.join(REF, "examples", "G2")
    for f in sorted(os.listdir(g2)):
        if f.endswith(".xyz"):
            open(os.path.join(HERE, "g2", f), "w").write(open(os.path.join(g2, f)).read())
    logs = os.path.join(REF, "docs", "validation", "sq")
    table = {}
    for f in sorted(os.listdir(logs)):
        m = re.match(r"(.+)_uhf_q(-?\d+)m(\d+)_(321g|631g_star)\.log$", f)
        if not m:
            continue
        txt = open(os.path.join(logs, f)).read()
        e = re.search(r"Done SCF Loop Total Energy is\s+(-?\d+\.\d+)", txt)
        c = re.search(r"Schwarz inequality screening cut-off\s+(\S+)", txt)
        rest = re.search(r"Requested (restricted|unrestricted)", txt)
        iters = re.findall(r"^\s+(\d+)\s+-?\d+\.\d+\s+\S+\s*$", txt.split("Enter SCF loop")[1].split("Done SCF")[0], re.M) if e else []
        if not e:
            continue
        table.setdefault(m.group(1), {})[m.group(4)] = {
            "energy": float(e.group(1)), "charge": int(m.group(2)), "mult": int(m.group(3)),
            "schwarz_cutoff": float(c.group(1)) if c else None,
            "restricted": rest.group(1) == "restricted" if rest else None,
            "iterations": int(iters[-1]) if iters else None,
            "source": "docs/validation/sq/" + f,
        }
    gms = os.path.join(REF, "docs", "validation", "gamess")
    for f in sorted(os.listdir(gms)):
        m = re.match(r"(.+)_uhf_631g_star\.log$", f)
        if not m or m.group(1) not in table:
            continue
        e = re.search(r"FINAL UHF ENERGY IS\s+(-?\d+\.\d+)", open(os.path.join(gms, f)).read())
        if e:  # first point of the optimisation = the G2 input geometry
            table[m.group(1)].setdefault("631g_star", {})["gamess_first_energy"] = float(e.group(1))
    json.dump(table, open(os.path.join(HERE, "g2_energies.json"), "w"), indent=1, sort_keys=True)
    print("wrote", len(table), "molecules")


if __name__ == "__main__":
    main()
