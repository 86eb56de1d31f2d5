"""The HIP solver against the oracle on the reference's own configurations: every case
whose host setup is pinned bit for bit by tests/test_setup_golden.py (the app's presets,
the staircase wall model, flight with a base body, the burst start, each injector element
type, injector hardware in the chamber), 60 steps from the reference's initial state,
injector cases with 4096 parcels.

Bars: relative L2 of rho, u, p <= 1e-4; v <= 1e-4 or twice the FP32 oracle's own distance
from the FP64 oracle at this horizon (v is near zero over much of these flows), and the
GPU's FP32 v no further from the FP64 oracle than that.
"""
import json
import os

import numpy as np
import pytest

import oracle as orc
from rnozzle._lib import lib
from rnozzle.sim_setup import SimSetup, default_config
from test_gpu_parity import fp32_noise, prims, rel_l2, run_pair, say

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
    noise = fp32_noise(s, STEPS, name, parcels=parcels)
    G = g.readAll()
    Ao, Bo = o.state()
    pg, po = prims(G["A"], G["B"]), prims(Ao, Bo)
    errs = {k: rel_l2(pg[k], po[k]) for k in pg}
    o64 = orc.Oracle(s.Nx, s.Nr, 64)
    o64.set_geometry(s.mask, s.geoA, s.geoG, s.geoM, s.jWall)
    if parcels:
        o64.set_parcels(parcels)
    o64.set_injector(s.injector)
    o64.set_state(*s.initial_state())
    o64.step(STEPS, s.P)
    p64 = prims(*o64.state_f64())
    e64 = {k: rel_l2(pg[k], p64[k]) for k in pg}
    say(f"[presets] {name} {s.Nx}x{s.Nr}: rel-L2 " + " ".join(f"{k}={v:.2e}" for k, v in errs.items())
        + " | vs FP64 oracle " + " ".join(f"{k}={v:.2e}" for k, v in e64.items()))
    assert np.isfinite(G["A"]).all() and np.isfinite(G["B"]).all()
    for k in ("rho", "u", "p"):
        assert errs[k] <= 1e-4, f"{name}: {k} rel-L2 {errs[k]:.3e}"
    bar = max(1e-4, 2 * noise["v"])
    assert errs["v"] <= bar and e64["v"] <= bar, (name, errs["v"], e64["v"], bar)
    np.testing.assert_allclose(g.readStepInfo(), o.step_info(), rtol=1e-5)
