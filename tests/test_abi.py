"""The C-ABI library loads and exports every function include/cuteflow_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

from cuteflow_amd import _lib
from cuteflow_amd.build import hdf5_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cf_\w+)\s*\(", text)))


def test_headers_are_covered():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == [
        "cuteflow_cgns.h", "cuteflow_hip.h"]


def test_library_exports_all_declared_symbols():
    lib = _lib.load()
    names = declared_functions("cuteflow_hip.h")
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes table covers exactly the declared API
    assert set(names) == set(_lib.SIGNATURES)


@pytest.mark.skipif(not hdf5_available(), reason="no HDF5 installation")
def test_cgns_library_exports_all_declared_symbols():
    lib = _lib.load_cgns()
    names = declared_functions("cuteflow_cgns.h")
    assert len(names) == 6
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.CGNS_SIGNATURES)
    assert ctypes.sizeof(_lib.CfCgnsInfo) == 8 * 4


def test_error_strings():
    lib = _lib.load()
    for s in (0, -1, -2, -3, -4, -5, -6, -7):
        assert lib.cf_error_string(s)
    assert lib.cf_error_string(-99) == b"unknown status"


def test_header_struct_sizes_match_ctypes():
    # cf_params: 8 doubles, 10 ints, 3*20 doubles, 20 ints, 2*20*64 doubles
    assert ctypes.sizeof(_lib.CfParams) == 8 * 8 + 10 * 4 + 60 * 8 + 20 * 4 + 2 * 20 * 64 * 8
    assert ctypes.sizeof(_lib.CfMesh) == 8 + 4 * 8
