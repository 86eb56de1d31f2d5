"""Integrated quantities on the device against the reference's own functions.

The fixtures hold, for each of the 13 configurations of tests/golden/setup_index.json,
the exit-plane column, the throat-plane column and the chamber cell of a seeded synthetic
flow, and what the reference's exitPerformance (simsetup.js:314-337), chamberPerformance
(239-250) and cellPrimSI (338-348) return for them (written by
tests/golden/make_setup_golden.js under node). The solver is given a state holding those
columns (setState), and nz_integrals_read must return the same numbers to 1e-12 relative:
both sum the same float32 texels in double, in a different order. The device sums are
also bit-stable from call to call, and the totals match a host sum over the fluid groups.
"""
import json
import os
import sys

import numpy as np
import pytest

from rnozzle._lib import lib
from rnozzle.sim_setup import SimSetup, default_config
from rnozzle.solver import GPUSolver, load_setup

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


def golden_state(s, g):
    """The initial state with the fixture's exit and throat columns and chamber cell."""
    A, B = (x.reshape(s.Nr, s.Nx, 4).copy() for x in s.initial_state())
    for i, a, b in ((s.iThroat, "throatA", "throatB"), (s.iExit, "exitA", "exitB")):
        A[:, i, :] = g[a].reshape(s.Nr, 4)
        B[:, i, :] = g[b].reshape(s.Nr, 4)
    A[0, s.iChamber, :] = g["cellA"]
    B[0, s.iChamber, :] = g["cellB"]
    return A.ravel(), B.ravel()


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("name", CASES)
def test_integrals_against_reference(name, prec):
    meta = INDEX[name]
    g = np.load(os.path.join(HERE, "golden", f"setup_{name}.npz"))
    s = SimSetup(default_config(**meta["cfg"]))
    gs = GPUSolver(precision=prec)
    load_setup(gs, s)
    A, B = golden_state(s, g)
    gs.setState(A, B)
    got = gs.integrals(s.P)
    ref = meta["integ"]
    pairs = [("F", ref["exit"]["F"]), ("mdot", ref["exit"]["mdot"]), ("Me", ref["exit"]["Me"]), ("pe", ref["exit"]["pe"]),
             ("Isp", ref["exit"]["Isp"]), ("mdotThroat", ref["chamber"]["mdot"]),
             ("unburntOx", ref["chamber"]["unburntOx"]), ("fuelVapour", ref["chamber"]["fuelVapour"]),
             ("Pc", ref["cell"]["p"]), ("Tc", ref["cell"]["T"])]
    worst = 0.0
    for k, want in pairs:
        e = abs(got[k] - want) / max(abs(want), 1e-300)
        worst = max(worst, e)
        assert got[k] == pytest.approx(want, rel=1e-12, abs=1e-14), (name, k, got[k], want)
    say(f"[integrals] {name} FP{prec}: 10 quantities against the reference's functions, worst rel {worst:.1e}")
    # the same bits on every call; totals = a host FP64 sum over fluid group hosts
    again = gs.integrals(s.P)
    assert all(again[k] == got[k] for k in got)
    st = gs.readAllF64() if prec == 64 else gs.readAll()
    Am = st["A"].reshape(-1, 4).astype(np.float64)
    M = s.geoM.reshape(-1, 4)
    host = (s.mask == 0) & (M[:, 0] == 0) & (M[:, 1] == 0)
    V = s.geoG.reshape(-1, 4)[:, 0].astype(np.float64)
    assert got["mass"] == pytest.approx(np.sum(Am[host, 0] * V[host]), rel=1e-13)
    assert got["energy"] == pytest.approx(np.sum(Am[host, 3] * V[host]), rel=1e-13)
