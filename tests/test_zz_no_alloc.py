"""nz_cutcell_build_noalloc (no device allocation, one static device array per GPU) against
nz_cutcell_build (device buffers allocated per call): bit-identical outputs on two calls in a
row, the second on a smaller grid (so the array holds the first call's larger contents),
through the C ABI; a grid beyond the array's size is refused with an error."""
import ctypes as C
import sys

import numpy as np
import pytest

from rnozzle import abi, configs
from rnozzle._lib import check, lib
from rnozzle.sim_setup import SimSetup

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(fn, s):
    inp = abi.NzCutcellIn()
    check(lib().nz_setup_get_cutcell_input(s._h, C.byref(inp)), "nz_setup_get_cutcell_input")
    N, NL = s.Nx * s.Nr, (s.Nx + 1) * (s.Nr + 1)
    out = {"mask": np.full(N, 7, np.uint8), "geoA": np.full(NL * 4, np.nan, np.float32),
           "geoG": np.full(N * 4, np.nan, np.float32), "geoM": np.full(N * 4, np.nan, np.float32),
           "kappa": np.full(N, np.nan, np.float32), "Ax": np.full(NL, np.nan, np.float32),
           "Ar": np.full(NL, np.nan, np.float32)}
    counts = (C.c_int32 * 2)(-1, -1)
    f = lambda k: out[k].ctypes.data_as(C.POINTER(C.c_float))
    rc = fn(C.byref(inp), out["mask"].ctypes.data_as(C.POINTER(C.c_uint8)), f("geoA"), f("geoG"), f("geoM"),
            f("kappa"), f("Ax"), f("Ar"), counts)
    return rc, out, (counts[0], counts[1])


def test_noalloc_matches_allocating_builder():
    if not lib().nz_device_available():
        pytest.fail("no gfx950 device: the HIP path must run on an MI355X")
    L = lib()
    # the largest batched grid, then a smaller injector geometry with chamber hardware
    cases = [("cone 1024x384", SimSetup(configs.cold_cone(Nx=1024))),
             ("injector hardware 192", SimSetup(configs.kerolox_injector(Nx=192)))]
    for name, s in cases:
        say(f"[noalloc] {name}: {s.Nx}x{s.Nr}")
        rc0, ref, c0 = build(L.nz_cutcell_build, s)
        check(rc0, "nz_cutcell_build")
        rc1, got, c1 = build(L.nz_cutcell_build_noalloc, s)
        check(rc1, "nz_cutcell_build_noalloc")
        for k in ref:
            a, b = ref[k], got[k]
            assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                  b.view(np.uint32) if b.dtype == np.float32 else b), (name, k)
        assert c0 == c1 and c1[0] > 0
        # and the host restatement's geometry
        assert np.array_equal(got["kappa"].view(np.uint32), s.kappa.view(np.uint32))
        say(f"[noalloc] {name}: bit-identical to nz_cutcell_build ({c1[0]} cut, {c1[1]} merged cells)")
    # beyond the array: refused before any device work
    big = SimSetup(configs.cold_cone(Nx=2048))
    rc, _, _ = build(L.nz_cutcell_build_noalloc, big)
    assert rc != 0
    msg = L.nz_last_error().decode()
    assert "device scratch" in msg, msg
    say(f"[noalloc] 2048x768 refused: {msg}")
