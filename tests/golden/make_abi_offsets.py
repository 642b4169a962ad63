"""Record the field offsets of the reference's ABI structs.

Compiles a small offsetof() probe against the reference headers
(src/basis.h, src/mol.h, src/option.h; non-LIBXC layout) and writes
tests/golden/abi_offsets.json.  Run in a container with the reference:
  python tests/golden/make_abi_offsets.py [/root/reference]
"""
import json
import os
import re
import subprocess
import sys
import tempfile

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def fields(struct, header):
    body = re.search(r"struct %s\s*\{(.*?)\n\};" % struct, open(header).read(), re.S).group(1)
    body = re.sub(r"#ifdef LIBXC.*?#endif", "", body, flags=re.S)
    body = re.sub(r"//.*", "", body)
    names = []
    for decl in body.split(";"):
        decl = "\n".join(l for l in decl.splitlines() if not l.strip().startswith("#")).strip()
        if not decl:
            continue
        parts = decl.replace("*", " ").split(None, 1)
        if len(parts) < 2:
            continue
        for nm in parts[1].split(","):
            names.append(re.sub(r"\[.*", "", nm.strip()))
    return names


def main():
    src = os.path.join(REF, "src")
    spec = {"GTOBasis_t": "basis.h", "Molecule_t": "mol.h", "option_t": "option.h"}
    prog = ['#include <stdio.h>', '#include <stddef.h>', '#include "basis.h"', '#include "mol.h"',
            '#include "option.h"', "int main(void){", 'printf("{");']
    first = True
    for st, h in spec.items():
        fl = fields(st, os.path.join(src, h))
        prog.append('printf("%s\\"%s\\": {\\"size\\": %%zu", sizeof(struct %s));' % ("" if first else ",", st, st))
        for f in fl:
            prog.append('printf(", \\"%s\\": %%zu", offsetof(struct %s, %s));' % (f, st, f))
        prog.append('printf("}");')
        first = False
    prog += ['printf("}\\n");', "return 0;}"]
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write("\n".join(prog))
        subprocess.run(["gcc", "-I", src, c, "-o", os.path.join(d, "p")], check=True)
        out = subprocess.run([os.path.join(d, "p")], check=True, capture_output=True, text=True).stdout
    data = json.loads(out)
    json.dump(data, open(os.path.join(HERE, "abi_offsets.json"), "w"), indent=1)
    print({k: v["size"] for k, v in data.items()})


if __name__ == "__main__":
    main()
