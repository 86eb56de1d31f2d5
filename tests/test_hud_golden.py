"""HUD read-outs (wallSeparation, summerfieldPoint; simsetup.js:352-384) in the C++ host
against the reference's own outputs on synthetic near-wall rows.

tests/golden/hud_golden.json was written by tests/golden/make_hud_golden.js, which runs
the reference's host JavaScript under node on three setups and eight kinds of rows each.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from rnozzle import abi
from rnozzle._lib import lib
from rnozzle.sim_setup import default_config, SimSetup

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "hud_golden.json")))


@pytest.mark.parametrize("name", sorted(GOLD))
def test_wall_separation_and_summerfield(name):
    case = GOLD[name]
    s = SimSetup(default_config(**case["cfg"]))
    found = 0
    for r in case["rows"]:
        row = np.asarray(r["row"], np.float32)
        out = abi.NzSeparation()
        got = lib().nz_wall_separation(s._h, row.ctypes.data_as(C.POINTER(C.c_float)), r["pa"], C.byref(out))
        if r["sep"] is None:
            assert got == 0, (name, r["kind"])
            continue
        found += 1
        assert got == 1, (name, r["kind"])
        for k in ("x", "frac", "eps", "pRatio", "Msep"):
            assert getattr(out, k) == pytest.approx(r["sep"][k], rel=1e-12, abs=1e-15), (name, r["kind"], k)
    assert found >= 2
    for sf in case["summerfield"]:
        frac = C.c_double()
        got = lib().nz_summerfield_point(s._h, sf["Pc"], sf["g"], sf["pa"], C.byref(frac))
        if sf["frac"] is None:
            assert got == 0
        else:
            assert got == 1 and frac.value == sf["frac"]


def test_wrappers_follow_the_reference_return_shape():
    """SimSetup.wall_separation / summerfield_point return what the reference's functions
    return: an object (dict) or null (None)."""
    name = sorted(GOLD)[0]
    case = GOLD[name]
    s = SimSetup(default_config(**case["cfg"]))
    for r in case["rows"]:
        got = s.wall_separation(np.asarray(r["row"], np.float32), r["pa"])
        if r["sep"] is None:
            assert got is None
        else:
            assert set(got) == {"x", "frac", "eps", "pRatio", "Msep"}
            assert got["frac"] == pytest.approx(r["sep"]["frac"], rel=1e-12, abs=1e-15)
    for sf in case["summerfield"]:
        got = s.summerfield_point(sf["Pc"], sf["g"], sf["pa"])
        assert got == sf["frac"]
    with pytest.raises(ValueError):
        s.wall_separation(np.zeros(3, np.float32), 101325.0)
