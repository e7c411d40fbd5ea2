"""The C-ABI library loads and exports every function include/cuteflow_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from cuteflow_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(cf_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_all_declared_symbols():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes table covers exactly the declared API
    assert set(names) == set(_lib.SIGNATURES)


def test_error_strings():
    lib = _lib.load()
    for s in (0, -1, -2, -3, -4, -5):
        assert lib.cf_error_string(s)
    assert lib.cf_error_string(-99) == b"unknown status"


def test_header_struct_sizes_match_ctypes():
    # cf_params: 8 doubles, 10 ints, 3*20 doubles, 20 ints, 2*20*64 doubles
    assert ctypes.sizeof(_lib.CfParams) == 8 * 8 + 10 * 4 + 60 * 8 + 20 * 4 + 2 * 20 * 64 * 8
    assert ctypes.sizeof(_lib.CfMesh) == 8 + 4 * 8
