"""Host setup (C++ restatement of simsetup.js / cutcell.js / physics.js) against the
reference's own outputs, bit for bit.

The fixtures in tests/golden/setup_*.npz were written by tests/golden/make_setup_golden.py,
which runs the reference's host JavaScript (buildSimSetup, computeParams,
buildCutGeometry, initialState, buildInjectorTable) under node on these configurations.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

from rnozzle import abi
from rnozzle._lib import lib
from rnozzle.sim_setup import SimSetup, default_config

HERE = os.path.dirname(os.path.abspath(__file__))
INDEX = json.load(open(os.path.join(HERE, "golden", "setup_index.json")))
CASES = sorted(k for k in INDEX if k not in ("atmosphere", "theory"))


def cfg_from_js(js):
    over = {}
    for k, v in js.items():
        over[k] = v
    return default_config(**over)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint8)


@pytest.fixture(scope="module", params=CASES)
def case(request):
    name = request.param
    meta = INDEX[name]
    g = np.load(os.path.join(HERE, "golden", f"setup_{name}.npz"))
    s = SimSetup(cfg_from_js(meta["cfg"]))
    return name, meta, g, s


def test_grid_and_probes(case):
    name, meta, g, s = case
    assert (s.Nx, s.Nr) == (meta["Nx"], meta["Nr"])
    assert s.dx == meta["dx"]
    assert (s.iExit, s.jExit, s.iChamber, s.iThroat) == (meta["iExit"], meta["jExit"], meta["iChamber"], meta["iThroat"])
    assert (s.info.nCut, s.info.nMerged) == (meta["nCut"], meta["nMerged"])
    for k in ("Rt", "Re", "Rc", "Lc", "xt", "xe", "eps", "thn", "the"):
        assert getattr(s.info, k) == pytest.approx(meta["contour"][k], rel=1e-14, abs=1e-15)


@pytest.mark.parametrize("field", ["mask", "geoA", "geoG", "geoM", "jWall", "kappa"])
def test_cut_geometry_bitwise(case, field):
    name, meta, g, s = case
    mine = getattr(s, field)
    ref = g[field]
    assert mine.shape == ref.shape
    assert np.array_equal(bits(mine), bits(ref)), f"{name}.{field}: {np.sum(bits(mine) != bits(ref))} values differ"


def test_initial_state_bitwise(case):
    name, meta, g, s = case
    A, B = s.initial_state()
    assert np.array_equal(bits(A), bits(g["A"]))
    assert np.array_equal(bits(B), bits(g["B"]))


def test_params(case):
    name, meta, g, s = case
    ref = g["P"]
    # double-precision P block equal to the reference's (libm may differ in the last ulp)
    np.testing.assert_allclose(s.P64, ref, rtol=1e-14, atol=1e-300)
    # and the float uniforms the GPU sees are identical
    assert np.array_equal(s.P64.astype(np.float32), ref.astype(np.float32))


def test_live_params(case):
    """computeParams with another chamber and start pressure (applyLive, app.js:285-304)."""
    name, meta, g, s = case
    Pc, pStart = meta["P2args"]
    P = s.compute_params(Pc, pStart)
    ref = g["P2"].astype(np.float32)
    mine = []
    for fname, n in abi.PARAM_ORDER:
        v = getattr(P, fname)
        mine.extend(list(v) if n > 1 else [v])
    assert np.array_equal(np.array(mine, np.float32), ref)


def test_injector_table(case):
    name, meta, g, s = case
    if not meta["injector"]:
        assert s.injector is None
        return
    assert np.array_equal(bits(s.injector), bits(g["tab"]))


def test_radius_at(case):
    name, meta, g, s = case
    xs, rs = g["radius_x"], g["radius"]
    mine = np.array([s.radius_at(x) for x in xs])
    np.testing.assert_allclose(mine, rs, rtol=1e-15, atol=0)


def test_atmosphere():
    out = (C.c_double * 4)()
    for h, ref in INDEX["atmosphere"]:
        lib().nz_atmosphere(h, out)
        assert out[0] == pytest.approx(ref["T"], rel=1e-15)
        assert out[1] == pytest.approx(ref["p"], rel=1e-14)
        assert out[2] == pytest.approx(ref["rho"], rel=1e-14)


def test_nozzle_theory():
    out = (C.c_double * 14)()
    names = ["At", "Ae", "eps", "Me", "pe", "Te", "ue", "cstar", "mdot", "F", "Fvac", "Isp", "c0", "uLim"]
    for args, ref in INDEX["theory"]:
        assert lib().nz_nozzle_theory(*args, out) == 0
        for k, n in enumerate(names):
            assert out[k] == pytest.approx(ref[n], rel=1e-13), n


def test_readme_ideal_numbers():
    """README.md:96: kerolox demo at 196 bar, sea level, 1-D ideal: 4.51 kg/s, 13.78 kN, 312 s."""
    out = (C.c_double * 14)()
    lib().nz_nozzle_theory(196e5, 3670, 1.22, 22, 0.023, 0.119, 101325, out)
    assert out[8] == pytest.approx(4.51, abs=0.006)
    assert out[9] / 1e3 == pytest.approx(13.78, abs=0.006)
    assert out[11] == pytest.approx(312, abs=0.6)
