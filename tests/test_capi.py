"""The C-ABI library loads and exports every symbol the public header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from aeroapps_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "aeroapps_amd.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|char|void)\s*\*?\s*(aa_\w+)\s*\(", txt, flags=re.M)))


def test_header_parsed():
    names = header_functions()
    assert "aa_vlidort_scalar_lambert" in names and "aa_get_aop_vector" in names
    assert len(names) >= 13


def test_library_exports_header_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} missing: build with python -m aeroapps_amd.build_lib")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_signatures_cover_header():
    assert set(header_functions()) == set(_lib.SIGNATURES)


def test_version_string():
    assert _lib.lib().aa_version().decode().startswith("aeroapps_amd")
