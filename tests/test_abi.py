"""The C-ABI library loads and exports every function include/pgfem3d_amd.h declares."""
import ctypes
import os
import re

import pgfem3d_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*int\s+(pgf_\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "pgf_create" in names and "pgf_residual" in names and "pgf_tangent" in names
    assert sorted(names) == sorted(P.exported_symbols())


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(P.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
