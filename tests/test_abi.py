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
    assert C.sizeof(abi.NzIntegrals) == 8 * 13


def test_last_error_and_null_handling():
    L = lib()
    assert L.nz_resize(None, 8, 8) != 0
    assert b"null" in L.nz_last_error()
    # the device probe never throws
    assert L.nz_device_available() in (0, 1)


def test_batch_structs():
    L = lib()
    assert L.nz_batch_instance_bytes(768, 256, 32, 0) > 768 * 256 * 64
    assert L.nz_batch_create(None, 32, 1 << 30, 4) != 0


def test_struct_layout_matches_header(tmp_path):
    """sizeof / offsetof of every struct the Python side mirrors, from the C compiler."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        import pytest
        pytest.skip("no C compiler")
    mirrors = {"nz_params": abi.NzParams, "nz_read_req": abi.NzReadReq, "nz_spray_stats_t": abi.NzSprayStats,
               "nz_probes": abi.NzProbes, "nz_integrals": abi.NzIntegrals, "nz_config": abi.NzConfig,
               "nz_setup_info": abi.NzSetupInfo, "nz_batch_job": abi.NzBatchJob, "nz_batch_result": abi.NzBatchResult,
               "nz_derived_params": abi.NzDerivedParams, "nz_separation": abi.NzSeparation,
               "nz_cutcell_in": abi.NzCutcellIn}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "nozzle.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
