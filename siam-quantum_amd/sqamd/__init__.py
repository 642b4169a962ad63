"""Python mirror of the MI355X two-electron engine (ctypes over lib/libsqamd.so).

The native library is the product; this module only marshals arrays.  It
raises if the library is not built: there is no Python or CPU fallback.
"""
from .abi import (Basis, GTOBasis, GTOBasisSet, Mol, Molecule, Option,  # noqa: F401
                  basis_from_ptr, default_option, dptr)
from .engine import JKEngine, lib, library_path  # noqa: F401
