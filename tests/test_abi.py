"""The C-ABI library loads and exports every entry point include/nozzle.h declares."""
import ctypes as C
import os
import re

from rnozzle import abi
from rnozzle._lib import LIB_PATH, lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "nozzle.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b(nz_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_built():
    assert os.path.exists(LIB_PATH)


def test_every_declared_symbol_exported():
    L = C.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert set(declared_functions()) <= set(abi.SIGNATURES)


def test_struct_sizes():
    # field-by-field mirrors of the header structs
    assert C.sizeof(abi.NzParams) == 4 * abi.N_PARAMS_F64 == 252
    assert C.sizeof(abi.NzReadReq) == 20
    assert C.sizeof(abi.NzIntegrals) == 8 * 11


def test_last_error_and_null_handling():
    L = lib()
    assert L.nz_resize(None, 8, 8) != 0
    assert b"null" in L.nz_last_error()
    # the device probe never throws
    assert L.nz_device_available() in (0, 1)
