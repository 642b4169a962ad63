"""The C-ABI library loads without a GPU and exports every symbol that
include/*.h declares; struct layouts match the reference headers."""
import ctypes as C
import os
import re

from sqamd import GTOBasis, Molecule, Option, library_path

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            if m.group(1) not in ("defined",):
                names.add(m.group(1))
    return names


def test_library_exports_declared_symbols():
    # RTLD_NOW: every symbol the library needs must resolve at load time
    lib = C.CDLL(library_path(), mode=os.RTLD_NOW)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_reference():
    """Offsets recorded from the reference headers (tests/golden/abi_offsets.json,
    written by make_abi_offsets.py) against the ctypes mirror and against
    include/sq_types.h compiled here."""
    import json
    import subprocess
    import tempfile
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "abi_offsets.json")))
    for cls, name in [(GTOBasis, "GTOBasis_t"), (Molecule, "Molecule_t"), (Option, "option_t")]:
        ref = gold[name]
        assert C.sizeof(cls) == ref["size"], name
        for f, _ in cls._fields_:
            assert getattr(cls, f).offset == ref[f], (name, f)
    prog = ['#include <stdio.h>', '#include <stddef.h>', '#include "sq_types.h"', "int main(void){"]
    for name, ref in gold.items():
        prog.append('printf("%%zu\\n", sizeof(struct %s));' % name)
        for f in ref:
            if f != "size":
                prog.append('printf("%%zu\\n", offsetof(struct %s, %s));' % (name, f))
    prog.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "p.c")
        open(c, "w").write("\n".join(prog))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", os.path.join(d, "p")], check=True)
        got = [int(x) for x in subprocess.run([os.path.join(d, "p")], capture_output=True, text=True,
                                              check=True).stdout.split()]
    want = []
    for name, ref in gold.items():
        want.append(ref["size"])
        want += [v for k, v in ref.items() if k != "size"]
    assert got == want


def test_class_launch_helpers_have_internal_linkage():
    """Each angular class is a separate translation unit of the same source
    with different -DSQ_L* values.  A launch helper template with external
    linkage would have the same mangled name and a different body in every
    one of them, and the linker would keep one class's copy for all classes
    (every class would run the (ss|ss) kernel).  Only the per-class entry
    points sqjk_launch_XXXX may be exported from the launcher code."""
    import shutil
    import subprocess
    import pytest
    if shutil.which("nm") is None:
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", library_path()],
                         capture_output=True, text=True, check=True).stdout
    leaked = [ln for ln in out.splitlines() if re.search(r"\blaunch_(mode|wave|from)\b", ln)]
    assert not leaked, leaked
