"""The HIP solver against the oracle on the reference's own configurations: every case
whose host setup is pinned bit for bit by tests/test_setup_golden.py (the app's presets,
the staircase wall model, flight with a base body, the burst start, each injector element
type, injector hardware in the chamber), 60 steps from the reference's initial state,
injector cases with 4096 parcels.

Bar (tests/parity_gate.py): relative L2 of rho, u, p and v's error over the velocity norm
<= 1e-4 against the FP32 oracle; against the FP64 oracle (printed) <= 1e-3.
"""
import json
import os

import numpy as np
import pytest

import oracle as orc
from rnozzle._lib import lib
from rnozzle.sim_setup import SimSetup, default_config
from parity_gate import gate
from test_gpu_parity import prims, run_pair

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INDEX = json.load(open(os.path.join(HERE, "golden", "setup_index.json")))
CASES = sorted(k for k in INDEX if k not in ("atmosphere", "theory"))
STEPS, PARCELS = 60, 4096


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not lib().nz_device_available():
        pytest.fail("no gfx950 device: the HIP path must run on an MI355X")


@pytest.mark.parametrize("name", CASES)
def test_reference_configuration(name):
    s = SimSetup(default_config(**INDEX[name]["cfg"]))
    parcels = PARCELS if s.injector is not None else None
    g, o = run_pair(s, STEPS, parcels=parcels)
    G = g.readAll()
    assert np.isfinite(G["A"]).all() and np.isfinite(G["B"]).all()
    pg = prims(G["A"], G["B"])
    gate(pg, prims(*o.state()), f"{name} {s.Nx}x{s.Nr}")
    o64 = orc.Oracle(s.Nx, s.Nr, 64)
    o64.set_geometry(s.mask, s.geoA, s.geoG, s.geoM, s.jWall)
    if parcels:
        o64.set_parcels(parcels)
    o64.set_injector(s.injector)
    o64.set_state(*s.initial_state())
    o64.step(STEPS, s.P)
    gate(pg, prims(*o64.state_f64()), f"{name}: GPU FP32 vs oracle FP64", bar=1e-3)
    np.testing.assert_allclose(g.readStepInfo(), o.step_info(), rtol=1e-5)
