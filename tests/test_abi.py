"""The C-ABI libraries load and export every function their headers declare.
No compute calls: this runs without a GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "solfec-1.0_amd", "lib")


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"^\s*#.*$", "", txt, flags=re.M)          # macros are not functions
    return sorted(set(re.findall(r"\b([A-Z][A-Za-z0-9]*_[A-Za-z0-9_]+)\s*\(", txt)) - {"SFGS_H", "SOLFEC_H"})


@pytest.mark.parametrize("lib,header", [("libsfgs.so", "sfgs.h"), ("libsolfec.so", "solfec.h")])
def test_library_exports_every_declared_symbol(lib, header):
    path = os.path.join(LIB, lib)
    assert os.path.exists(path), f"{path} not built"
    if lib == "libsolfec.so":
        ctypes.CDLL(os.path.join(LIB, "libsfgs.so"))
    L = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_oracle_exports():
    L = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    txt = open(os.path.join(ROOT, "oracle", "oracle.h")).read()
    names = set(re.findall(r"\b(OR_[A-Za-z0-9_]+)\s*\(", txt))
    assert names and all(hasattr(L, n) for n in names)


def test_error_strings_without_gpu():
    L = ctypes.CDLL(os.path.join(LIB, "libsfgs.so"))
    L.SFGS_Error_String.restype = ctypes.c_char_p
    assert L.SFGS_Error_String(0) == b"ok"
    assert L.SFGS_Error_String(-2) == b"invalid argument"
