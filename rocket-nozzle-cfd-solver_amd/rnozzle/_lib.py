"""Loader for the native solver library (librnozzle.so, built from csrc/).

The library is built in-tree by `rnozzle.build.build()`; nothing here falls back
to Python: a missing or stale library raises at load time.
"""
import ctypes as C
import os

from . import abi

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(ROOT, "build_out", "librnozzle.so")

_lib = None


class NozzleError(RuntimeError):
    """Raised when a C ABI call returns a non-zero status (the reference throws Error)."""


def lib():
    """The loaded library with every header signature bound (raises if absent)."""
    global _lib
    if _lib is None:
        path = LIB_PATH
        if not os.path.exists(path):
            raise NozzleError(f"native library not built: {path} (run __graft_entry__.build())")
        L = C.CDLL(path)
        for name, (res, args) in abi.SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().nz_last_error().decode(errors="replace")
        raise NozzleError(f"{what}: {msg} (code {rc})")
    return rc
