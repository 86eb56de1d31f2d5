"""The batched and serving workloads on the GPU, each solve run to its own stop.

- Batched set (a seeded subset of the 64-solve set: its smallest and largest grid sizes,
  128 x 48 and 1024 x 384, and three drawn sizes): every solve runs to a residual of 1e-6 or
  3 ms of simulated time in the shared launches. Each solve's exit-plane thrust and mass
  flow, throat mass flow and chamber pressure equal those of the same solve run alone for the
  same number of steps within 0.1 % (they are bit-identical: the state is compared too).
- Serving: requests submitted between run calls while others are resident: nozzles of the
  batched set's distribution and 384 x 128 spray injectors with 16,384 parcels of each
  element type (showerhead, coaxial, impinging, pintle). Each result equals the same solve
  alone, state and integrated quantities.
"""
import sys

import numpy as np
import pytest

from rnozzle import abi, configs
from rnozzle._lib import lib
from rnozzle.batch import Batch
from rnozzle.sim_setup import SimSetup
from rnozzle.solver import GPUSolver, load_setup

pytestmark = pytest.mark.gpu
KEYS = ("F", "mdot", "mdotThroat", "Pc")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not lib().nz_device_available():
        pytest.fail("no gfx950 device: the HIP path must run on an MI355X")


def alone(setup, steps, parcels=None):
    g = GPUSolver()
    if parcels:
        g.resize(setup.Nx, setup.Nr)
        g.spray.setCount(parcels)
    load_setup(g, setup)
    left = steps
    while left > 0:
        n = min(left, 1000)
        g.step(n, setup.P)
        left -= n
    st = g.readAll()
    return st["A"], st["B"], g.integrals(setup.P)


def check_against_alone(b, jid, s, name, parcels=None):
    res = b.result(jid)
    assert res["state"] == abi.NZ_JOB_DONE and res["have_integrals"]
    A, B = b.read_state(jid)
    A1, B1, I1 = alone(s, res["steps"], parcels)
    same = np.array_equal(A, A1) and np.array_equal(B, B1)
    rel = {k: abs(res["integ"][k] / I1[k] - 1) if I1[k] != 0 else abs(res["integ"][k]) for k in KEYS}
    stop = {abi.NZ_STOP_RESIDUAL: "residual", abi.NZ_STOP_TIME: "time", abi.NZ_STOP_STEPS: "steps"}[res["stop"]]
    say(f"[workload] {name} {s.Nx}x{s.Nr}: {res['steps']} steps, stop {stop}, t = {res['t_grid'] * s.tScale * 1e3:.3f} ms; "
        f"F {res['integ']['F']:.4g} N, mdot {res['integ']['mdot']:.4g} kg/s, Pc {res['integ']['Pc'] / 1e5:.4g} bar; "
        f"alone: bit-identical={same}, worst integral {max(rel.values()):.1e}")
    assert same, name
    for k, v in rel.items():
        assert v < 1e-3, (name, k, v)
    return res


def test_batched_set_to_stops():
    rng = np.random.default_rng(20240501)
    setups = [("128x48", SimSetup(configs.batch_instance(rng, Nx=128))),
              ("1024x384", SimSetup(configs.batch_instance(rng, Nx=1024)))]
    setups += [(f"drawn {k}", SimSetup(configs.batch_instance(rng))) for k in range(3)]
    b = Batch(max_resident=16)
    ids = []
    for name, s in setups:
        ids.append(b.submit(s, t_end_s=configs.BATCH_T_END, res_tol=configs.BATCH_RES_TOL, min_steps=100,
                            keep_state=True))
    runs = 0
    while True:
        r, q = b.run(200)
        runs += 1
        if runs % 25 == 0:
            say(f"[workload] batched set: {r} resident after {runs} runs")
        if r == 0 and q == 0:
            break
    for (name, s), jid in zip(setups, ids):
        res = check_against_alone(b, jid, s, f"batched {name}")
        assert res["stop"] in (abi.NZ_STOP_RESIDUAL, abi.NZ_STOP_TIME)
        if res["stop"] == abi.NZ_STOP_TIME:
            assert res["t_grid"] * s.tScale >= configs.BATCH_T_END * (1 - 1e-6)


def test_serving_midflight_arrivals():
    rng = np.random.default_rng(99)
    b = Batch(max_resident=16)
    first = [("nozzle a", SimSetup(configs.batch_instance(rng, Nx=256)), None),
             ("nozzle b", SimSetup(configs.batch_instance(rng, Nx=512)), None)]
    later = [(f"injector {k}", SimSetup(configs.injector_instance(rng, kind=k)), configs.INJECTOR_PARCELS)
             for k in ("shower", "coax", "impinge", "pintle")]
    later.append(("nozzle c", SimSetup(configs.batch_instance(rng, Nx=128)), None))
    submitted = []
    for name, s, parcels in first:
        submitted.append((name, s, parcels, b.submit(s, t_end_s=configs.BATCH_T_END, res_tol=configs.BATCH_RES_TOL,
                                                     min_steps=100, keep_state=True, parcels=parcels)))
    for _ in range(5):        # the first solves are resident and running before the others arrive
        r, q = b.run(200)
        assert r == 2
    for name, s, parcels in later:
        submitted.append((name, s, parcels, b.submit(s, t_end_s=configs.BATCH_T_END, res_tol=configs.BATCH_RES_TOL,
                                                     min_steps=100, keep_state=True, parcels=parcels)))
        r, q = b.run(100)     # each arrives while the others run
        assert r >= 2
    while True:
        r, q = b.run(200)
        if r == 0 and q == 0:
            break
    for name, s, parcels, jid in submitted:
        res = check_against_alone(b, jid, s, f"served {name}", parcels)
        if parcels:
            assert s.injector is not None
            assert res["t_admit"] > b.result(submitted[0][3])["t_admit"]
