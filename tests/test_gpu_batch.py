"""Batched solves (nz_batch_*): each solve's result is the same alone and in the mix.

"Alone" is a GPUSolver on the same inputs stepped the same number of steps. Every solve
uses the same kernels and the same per-instance arithmetic, and the spray deposit sums
fixed-point integers (whose order of addition leaves no trace), so the bar is bit-identity
of every state channel, with and without spray.
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not lib().nz_device_available():
        pytest.fail("no gfx950 device: the HIP path must run on an MI355X")


def rel_l2(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def alone(setup, steps, parcels=None):
    g = GPUSolver()
    if parcels:
        g.resize(setup.Nx, setup.Nr)
        g.spray.setCount(parcels)
    load_setup(g, setup)
    g.step(steps, setup.P)
    out = g.readAll()
    return g, out["A"], out["B"]


def mix_setups():
    rng = np.random.default_rng(7)
    return [
        ("cold cone 256", SimSetup(configs.cold_cone()), None),
        ("hot bell 192", SimSetup(configs.hot_bell(Nx=192)), None),
        ("inviscid cone 256", SimSetup(configs.apply(configs.cold_cone(), {"viscous": 0})), None),
        ("batch instance 128", SimSetup(configs.batch_instance(rng, Nx=128)), None),
        ("batch instance 320", SimSetup(configs.batch_instance(rng, Nx=320)), None),
        ("kerolox injector 160", SimSetup(configs.kerolox_injector(Nx=160)), 4096),
    ]


def run_to_end(b, chunk=25, limit=100000):
    done = []
    for _ in range(limit):
        r, q = b.run(chunk)
        done += b.finished()
        if r == 0 and q == 0:
            return done
    raise AssertionError("batch did not drain")


def test_mix_matches_alone():
    steps = 60
    setups = mix_setups()
    b = Batch(max_resident=16)
    ids = []
    for name, s, parcels in setups:
        say(f"[batch] submit {name} {s.Nx}x{s.Nr} parcels={parcels}")
        ids.append(b.submit(s, max_steps=steps, keep_state=True, parcels=parcels))
    done = run_to_end(b, chunk=25)
    assert sorted(done) == sorted(ids)
    for (name, s, parcels), jid in zip(setups, ids):
        res = b.result(jid)
        assert res["state"] == abi.NZ_JOB_DONE and res["stop"] == abi.NZ_STOP_STEPS and res["steps"] == steps
        A, B = b.read_state(jid)
        g, A1, B1 = alone(s, steps, parcels)
        same = np.array_equal(A, A1) and np.array_equal(B, B1)
        eA, eB = rel_l2(A, A1), rel_l2(B, B1)
        say(f"[batch] {name}: in the mix vs alone rel-L2 A={eA:.2e} B={eB:.2e} bit-identical={same}")
        assert same, name     # the spray deposit too: its sums are fixed-point integers
        # step info of the last step agrees with the solver's
        si = g.readStepInfo()
        assert abs(res["t_grid"] - float(si[2])) <= 1e-5 * max(1.0, abs(float(si[2])))
        # integrated quantities at retirement match the solver's device sums
        ig = g.integrals(s.P)
        for k in ("F", "mdot", "Pc", "mass"):
            assert res["integ"][k] == pytest.approx(ig[k], rel=1e-4, abs=1e-9), (name, k)
        b.release(jid)


def test_batch_with_a_large_instance():
    """A batch holding an instance above 2048 tiles closes the step with its own kernel
    after every step, for every resident; each result stays bit-identical to its solve
    run alone (where the large one also closes every step and the small one folds the
    step inside its next stage 1: the fold is the same fixed-order sum either way)."""
    steps = 30
    setups = [("cold cone 2048", SimSetup(configs.cold_cone(Nx=2048))),
              ("batch instance 128", SimSetup(configs.batch_instance(np.random.default_rng(3), Nx=128)))]
    assert (setups[0][1].Nx // 64) * (setups[0][1].Nr // 4) > 2048
    b = Batch(max_resident=4)
    ids = [b.submit(s, max_steps=steps, keep_state=True) for _, s in setups]
    done = run_to_end(b, chunk=10)
    assert sorted(done) == sorted(ids)
    for (name, s), jid in zip(setups, ids):
        assert b.result(jid)["steps"] == steps
        A, B = b.read_state(jid)
        _, A1, B1 = alone(s, steps)
        same = np.array_equal(A, A1) and np.array_equal(B, B1)
        say(f"[batch] {name} with a 6144-tile neighbour: bit-identical to alone {same}")
        assert same
        b.release(jid)


def test_graph_and_plain_launches_agree():
    s = SimSetup(configs.hot_bell(Nx=192))
    out = []
    for graph in (True, False):
        b = Batch(max_resident=4, graph=graph)
        jid = b.submit(s, max_steps=50, keep_state=True)
        run_to_end(b, chunk=10)
        out.append(b.read_state(jid))
        b.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_admission_by_budget_and_time_stop():
    rng = np.random.default_rng(11)
    setups = [SimSetup(configs.batch_instance(rng, Nx=nx)) for nx in (256, 192, 320, 128, 256)]
    need = [Batch.instance_bytes(s.Nx, s.Nr) for s in setups]
    budget = max(need) + min(need)  # at most two residents at a time
    b = Batch(budget_bytes=budget, max_resident=8)
    ids = [b.submit(s, t_end_s=2e-5, keep_state=False) for s in setups]
    peak, seen_queue = 0, False
    done = []
    for _ in range(100000):
        r, q = b.run(20)
        used, bud = b.memory()
        assert used <= bud
        peak = max(peak, r)
        seen_queue |= q > 0
        done += b.finished()
        if r == 0 and q == 0:
            break
    say(f"[batch] budget {budget} B: peak resident {peak}, queued seen {seen_queue}")
    assert sorted(done) == sorted(ids) and seen_queue and peak <= 2
    for s, jid in zip(setups, ids):
        res = b.result(jid)
        assert res["stop"] == abi.NZ_STOP_TIME
        tEnd = float(np.float32(2e-5 / s.tScale))  # the device holds the stop time as float
        assert res["t_grid"] >= tEnd
        # the same solve alone reaches the time in the same number of steps
        g = GPUSolver()
        load_setup(g, s)
        n = 0
        while float(g.readStepInfo()[2]) < tEnd:
            g.step(1, s.P)
            n += 1
            assert n < 100000
        assert n == res["steps"], (n, res["steps"])
        assert res["t_admit"] >= res["t_submit"] and res["t_done"] >= res["t_admit"]


def test_slab_reuse_and_pipelined_run():
    """Every resident's buffers come from the batch's one device slab; with room for two
    solves at a time, later admissions take the memory and state slots of retired ones,
    and each result is still bit-identical to the solve alone. A run call enqueues its
    steps before it reads the previous call's stop flags: the results are filed a call
    later, run(0) only admits (the last filed list stays readable) and the device time of
    the chunks read is positive and below the wall time of the loop."""
    import time
    rng = np.random.default_rng(5)
    setups = [SimSetup(configs.batch_instance(rng, Nx=nx)) for nx in (192, 128, 256, 128, 192)]
    need = [Batch.instance_bytes(s.Nx, s.Nr) for s in setups]
    b = Batch(budget_bytes=2 * max(need), max_resident=8)
    steps = [40, 25, 30, 55, 35]
    ids = [b.submit(s, max_steps=n, keep_state=True) for s, n in zip(setups, steps)]
    done, peak = [], 0
    t0 = time.perf_counter()
    for _ in range(100000):
        r, q = b.run(10)
        used, bud = b.memory()
        assert used <= bud
        fin = b.finished()
        done += fin
        peak = max(peak, r)
        if fin:
            r0, q0 = b.run(0)            # admission only: nothing filed, the list unchanged
            assert b.finished() == fin and r0 + q0 == r + q
        if r == 0 and q == 0:
            break
    wall = time.perf_counter() - t0
    busy = b.gpu_busy()
    say(f"[batch] slab reuse: peak resident {peak}, busy {busy:.3f} s of {wall:.3f} s")
    assert sorted(done) == sorted(ids) and peak < len(ids)
    assert 0 < busy <= wall
    for s, n, jid in zip(setups, steps, ids):
        res = b.result(jid)
        assert res["state"] == abi.NZ_JOB_DONE and res["steps"] == n
        A, B = b.read_state(jid)
        _, A1, B1 = alone(s, n)
        assert np.array_equal(A, A1) and np.array_equal(B, B1), (s.Nx, n)


def test_rejects_oversized_job():
    s = SimSetup(configs.cold_cone())
    b = Batch(budget_bytes=1 << 20, max_resident=2)
    with pytest.raises(Exception):
        b.submit(s, max_steps=10)


def test_batch_autotune_keeps_results():
    """The live-batch autotuner advances the solves with every tile shape in turn; the
    results stay bit-identical to the solves run alone (every shape computes the same)."""
    rng = np.random.default_rng(3)
    setups = [SimSetup(configs.batch_instance(rng, Nx=nx)) for nx in (256, 128)] + [SimSetup(configs.hot_bell(Nx=192))]
    steps = 100
    b = Batch(max_resident=8)
    ids = [b.submit(s, max_steps=steps, keep_state=True) for s in setups]
    best, us = b.autotune(4)
    say(f"[batch] autotune: best {best}, us/step " + " ".join(f"{u:.1f}" for u in us))
    assert all(u > 0 for u in us) and us[best] == min(us)
    run_to_end(b, chunk=16)
    for s, jid in zip(setups, ids):
        assert b.result(jid)["steps"] == steps
        A, B = b.read_state(jid)
        _, A1, B1 = alone(s, steps)
        assert np.array_equal(A, A1) and np.array_equal(B, B1)
