"""The cut-cell geometry built by the GPU kernels (cutcell.hip, buildCutGeometry of
cutcell.js:14-186) against the host restatement and the reference's own output.

For each of the 13 configurations whose setup fixtures the reference's JavaScript wrote
(tests/golden/setup_*.npz): the whole setup built with the geometry on the device
(nz_config.deviceGeometry = 1) gives mask, geoA, geoG, geoM, jWall and kappa bit for bit
equal to the fixtures and to the host-built setup; the builder called directly gives the
host's face areas, kappa and cut / merged counts bit for bit.
"""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest

from rnozzle import abi
from rnozzle._lib import check, lib
from rnozzle.sim_setup import SimSetup, default_config

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INDEX = json.load(open(os.path.join(HERE, "golden", "setup_index.json")))
CASES = sorted(k for k in INDEX if k not in ("atmosphere", "theory"))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not lib().nz_device_available():
        pytest.fail("no gfx950 device: the HIP path must run on an MI355X")


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def device_build(s):
    """nz_cutcell_build on the setup's own input; host arrays out."""
    inp = abi.NzCutcellIn()
    check(lib().nz_setup_get_cutcell_input(s._h, C.byref(inp)), "nz_setup_get_cutcell_input")
    N, NL = s.Nx * s.Nr, (s.Nx + 1) * (s.Nr + 1)
    out = {"mask": np.zeros(N, np.uint8), "geoA": np.zeros(NL * 4, np.float32), "geoG": np.zeros(N * 4, np.float32),
           "geoM": np.zeros(N * 4, np.float32), "kappa": np.zeros(N, np.float32), "Ax": np.zeros(NL, np.float32),
           "Ar": np.zeros(NL, np.float32)}
    counts = (C.c_int32 * 2)()
    f = lambda k: out[k].ctypes.data_as(C.POINTER(C.c_float))
    check(lib().nz_cutcell_build(C.byref(inp), out["mask"].ctypes.data_as(C.POINTER(C.c_uint8)), f("geoA"), f("geoG"),
                                 f("geoM"), f("kappa"), f("Ax"), f("Ar"), counts), "nz_cutcell_build")
    return out, (counts[0], counts[1])


@pytest.mark.parametrize("name", CASES)
def test_device_geometry_bitwise(name):
    meta = INDEX[name]
    g = np.load(os.path.join(HERE, "golden", f"setup_{name}.npz"))
    host = SimSetup(default_config(**meta["cfg"]))
    dev = SimSetup(default_config(**meta["cfg"], deviceGeometry=1))
    assert (dev.Nx, dev.Nr) == (host.Nx, host.Nr)
    for field in ("mask", "geoA", "geoG", "geoM", "jWall", "kappa"):
        mine = getattr(dev, field)
        assert np.array_equal(bits(mine), bits(getattr(host, field))), f"{name}.{field} differs from the host build"
        assert np.array_equal(bits(mine), bits(g[field])), f"{name}.{field} differs from the reference"
    assert (dev.info.nCut, dev.info.nMerged) == (meta["nCut"], meta["nMerged"])
    # the builder on its own: face areas, kappa, counts (geoM before the setup's eddy-length
    # and boundary-layer columns are written into it)
    out, counts = device_build(host)
    assert np.array_equal(bits(out["Ax"]), bits(host.Ax)) and np.array_equal(bits(out["Ar"]), bits(host.Ar))
    assert np.array_equal(bits(out["kappa"]), bits(host.kappa))
    assert np.array_equal(out["mask"], host.mask)
    assert np.array_equal(bits(out["geoA"]), bits(host.geoA)) and np.array_equal(bits(out["geoG"]), bits(host.geoG))
    M, Mh = out["geoM"].reshape(-1, 4), host.geoM.reshape(-1, 4)
    assert np.array_equal(bits(M[:, :2]), bits(Mh[:, :2]))
    assert np.array_equal(M[:, 2], Mh[:, 2] % 16)      # slave bits; the setup adds 16 in the boundary-layer band
    assert counts == (meta["nCut"], meta["nMerged"])
    say(f"[cutcell] {name} {host.Nx}x{host.Nr}: device geometry bit-identical "
        f"({counts[0]} cut, {counts[1]} merged cells)")
