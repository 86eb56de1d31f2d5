"""HIP solver against the CPU oracle (a pass-by-pass restatement of the reference's
GLSL), called through the C ABI on the same seeded inputs.

Tolerances (written here, per field):
- FP32 GPU vs FP32 oracle after a short horizon: relative L2 of rho, u, p <= 1e-4, and
  v's error over the velocity norm |(u, v)| <= 1e-4 (tests/parity_gate.py: v is near zero
  over most of these flows, where its own relative L2 is rounding noise; it is printed
  beside); step info to 1e-5.
- FP64 GPU vs FP64 oracle: relative L2 <= 1e-9 (same arithmetic up to operation
  order), which pins the kernel logic independently of FP32 rounding.
The error of every FP32 result against the FP64 run is printed.
"""
import sys

import numpy as np
import pytest

from rnozzle import configs
from rnozzle._lib import lib
from rnozzle.sim_setup import SimSetup
from rnozzle.solver import GPUSolver, load_setup

pytestmark = pytest.mark.gpu

import oracle as orc  # noqa: E402
from parity_gate import gate  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not lib().nz_device_available():
        pytest.fail("no gfx950 device: the HIP path must run on an MI355X")


def prims(A, B):
    A = A.reshape(-1, 4).astype(np.float64)
    B = B.reshape(-1, 4).astype(np.float64)
    rho = np.maximum(A[:, 0], 1e-7)
    u, v = A[:, 1] / rho, A[:, 2] / rho
    G = np.maximum(B[:, 2], 0.5)
    p = np.maximum((A[:, 3] - 0.5 * rho * (u * u + v * v)) / G, 1e-7)
    return {"rho": rho, "u": u, "v": v, "p": p}


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_pair(setup, steps, prec=32, parcels=None, chunks=1):
    """Advance the GPU solver and the oracle from the same inputs."""
    say(f"[parity] Nx={setup.Nx} Nr={setup.Nr} steps={steps} prec={prec} parcels={parcels}")
    g = GPUSolver(precision=prec)
    if parcels:
        g.resize(setup.Nx, setup.Nr)
        g.spray.setCount(parcels)
    A0, B0 = load_setup(g, setup) if not parcels else (None, None)
    if parcels:
        g.setGeometry({"mask": setup.mask, "geoA": setup.geoA, "geoG": setup.geoG, "geoM": setup.geoM, "jWall": setup.jWall})
        g.spray.setInjector({"tab": setup.injector})
        A0, B0 = setup.initial_state()
        g.setState(A0, B0)
    o = orc.Oracle(setup.Nx, setup.Nr, prec)
    o.set_geometry(setup.mask, setup.geoA, setup.geoG, setup.geoM, setup.jWall)
    if parcels:
        o.set_parcels(parcels)
    o.set_injector(setup.injector)
    o.set_state(A0, B0)
    for _ in range(chunks):
        g.step(steps // chunks, setup.P)
        o.step(steps // chunks, setup.P)
    g.sync()
    return g, o


def compare_state(g, o, tol, label):
    if g.precision == 64:
        G = g.readAllF64()
        Ao, Bo = o.state_f64()
    else:
        G = g.readAll()
        Ao, Bo = o.state()
    pg, po = prims(G["A"], G["B"]), prims(Ao, Bo)
    assert np.isfinite(G["A"]).all() and np.isfinite(G["B"]).all()
    if g.precision == 64:
        errs = {k: rel_l2(pg[k], po[k]) for k in pg}
        say(f"[parity] {label}: rel-L2 " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
        for k, v in errs.items():
            assert v <= tol, f"{label}: {k} rel-L2 {v:.3e} > {tol}"
    else:
        errs = gate(pg, po, label, tol)
    si_g, si_o = g.readStepInfo(), o.step_info()
    np.testing.assert_allclose(si_g, si_o, rtol=max(tol, 1e-5))
    return errs


@pytest.fixture(scope="module")
def cold():
    return SimSetup(configs.cold_cone())


def test_cold_cone_fp32(cold):
    g, o = run_pair(cold, 200)
    compare_state(g, o, 1e-4, "cold cone FP32 200 steps")
    assert g.stepCount == 200


def test_cold_cone_fp64_logic(cold):
    g, o = run_pair(cold, 100, prec=64)
    compare_state(g, o, 1e-9, "cold cone FP64 100 steps")


def test_cold_cone_fp32_vs_fp64(cold):
    """Error of the FP32 GPU result against the FP64 GPU run of the same kernels."""
    g32, _ = run_pair(cold, 200)
    g64 = GPUSolver(precision=64)
    load_setup(g64, cold)
    g64.step(200, cold.P)
    a, b = g32.readAll(), g64.readAll()
    p32, p64 = prims(a["A"], a["B"]), prims(b["A"], b["B"])
    errs = {k: rel_l2(p32[k], p64[k]) for k in p32}
    say("[parity] cold cone FP32 vs FP64 (200 steps): " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    assert max(errs.values()) < 1e-3


def fp32_noise(setup, steps, label, parcels=None):
    """Relative L2 of the FP32 oracle against the FP64 oracle: the rounding noise any
    FP32 implementation carries at this horizon."""
    o32 = orc.Oracle(setup.Nx, setup.Nr, 32)
    o64 = orc.Oracle(setup.Nx, setup.Nr, 64)
    A0, B0 = setup.initial_state()
    for o in (o32, o64):
        o.set_geometry(setup.mask, setup.geoA, setup.geoG, setup.geoM, setup.jWall)
        if parcels:
            o.set_parcels(parcels)
        o.set_injector(setup.injector)
        o.set_state(A0, B0)
        o.step(steps, setup.P)
    a, b = o32.state_f64(), o64.state_f64()
    p32, p64 = prims(a[0], a[1]), prims(b[0], b[1])
    errs = {k: rel_l2(p32[k], p64[k]) for k in p32}
    say(f"[parity] {label}: oracle FP32 vs FP64 " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    return errs


def test_inviscid_cold_cone():
    s = SimSetup(configs.apply(configs.cold_cone(), {"viscous": 0}))
    assert not s.P.visc and not s.P.chem
    g, o = run_pair(s, 100)
    G = g.readAll()
    gate(prims(G["A"], G["B"]), prims(*o.state()), "inviscid cold cone")


def test_inviscid_fp64_logic():
    s = SimSetup(configs.apply(configs.cold_cone(), {"viscous": 0}))
    g, o = run_pair(s, 150, prec=64)
    compare_state(g, o, 1e-9, "inviscid cold cone FP64")


def test_hot_bell_full_physics():
    """The hot bell at 192 x 64 with the full viscous / wall / afterburning split step,
    120 steps. (The ambient region above the chamber at the left boundary carries a mode
    that grows from rounding level, about x40 per 16 steps in the FP64 oracle, which puts
    v's plain relative L2 at 1e-3 between any two FP32 rounding sequences; measured against
    the velocity norm it is 4e-5.) The GPU FP32 result is also printed against the FP64
    oracle."""
    s = SimSetup(configs.hot_bell(Nx=192))
    assert s.P.visc and s.P.chem
    g, o = run_pair(s, 120, chunks=2)
    G = g.readAll()
    pg = prims(G["A"], G["B"])
    gate(pg, prims(*o.state()), "hot bell 192 full physics")
    o64 = orc.Oracle(s.Nx, s.Nr, 64)
    o64.set_geometry(s.mask, s.geoA, s.geoG, s.geoM, s.jWall)
    o64.set_state(*s.initial_state())
    o64.step(120, s.P)
    gate(pg, prims(*o64.state_f64()), "hot bell 192 full physics, GPU FP32 vs oracle FP64", bar=1e-3)
    np.testing.assert_allclose(g.readStepInfo(), o.step_info(), rtol=1e-5)


@pytest.mark.parametrize("name,parcels", [("hot_bell", None), ("kerolox_injector", 65536)])
def test_full_size_workload(name, parcels):
    """The two 768x256 workloads at their full size (the hot bell with full physics, the
    kerolox injector with 65,536 parcels), 120 steps, against the FP32 oracle at the fixed
    bar; the GPU's FP32 result against the FP64 oracle is printed (and held to 1e-3)."""
    s = SimSetup(getattr(configs, name)())
    assert (s.Nx, s.Nr) == (768, 256)
    g, o = run_pair(s, 120, parcels=parcels)
    G = g.readAll()
    pg = prims(G["A"], G["B"])
    gate(pg, prims(*o.state()), f"{name} 768x256")
    o64 = orc.Oracle(s.Nx, s.Nr, 64)
    o64.set_geometry(s.mask, s.geoA, s.geoG, s.geoM, s.jWall)
    if parcels:
        o64.set_parcels(parcels)
    o64.set_injector(s.injector)
    o64.set_state(*s.initial_state())
    o64.step(120, s.P)
    gate(pg, prims(*o64.state_f64()), f"{name} 768x256, GPU FP32 vs oracle FP64", bar=1e-3)
    np.testing.assert_allclose(g.readStepInfo(), o.step_info(), rtol=1e-5)


@pytest.mark.parametrize("name,parcels,tol", [("hot_bell", None, 1e-9), ("kerolox_injector", 65536, 1e-6)])
def test_full_size_fp64_logic(name, parcels, tol):
    """The same workloads in FP64 against the FP64 oracle, 40 steps: the fused kernels'
    logic at full size (parcels are stored as float32 on the device in both precisions,
    hence the injector's looser bar)."""
    s = SimSetup(getattr(configs, name)())
    g, o = run_pair(s, 40, prec=64, parcels=parcels)
    compare_state(g, o, tol, f"{name} 768x256 FP64")


def test_large_grid_closes_every_step():
    """Above 2048 tiles the step is closed by its own kernel after every step instead of
    inside the next stage 1: the cold cone at 2048 x 768 (6144 tiles of 64 x 4), 150 steps
    (by then the start-up has reached beyond the inlet corner; at 40 steps the whole
    velocity field is rounding-level), against the oracle at the fixed bar."""
    s = SimSetup(configs.cold_cone(Nx=2048))
    assert (s.Nx // 64) * (s.Nr // 4) > 2048
    g, o = run_pair(s, 150)
    assert g.tileShape() == 0
    G = g.readAll()
    gate(prims(G["A"], G["B"]), prims(*o.state()), "cold cone 2048 (per-step close) 150 steps")
    np.testing.assert_allclose(g.readStepInfo(), o.step_info(), rtol=1e-5)


def test_hot_bell_fp64_logic():
    s = SimSetup(configs.hot_bell(Nx=192))
    g, o = run_pair(s, 80, prec=64)
    compare_state(g, o, 1e-9, "hot bell 192 FP64")


def test_injector_spray():
    s = SimSetup(configs.kerolox_injector(Nx=160))
    assert s.P.inj and s.injector is not None
    g, o = run_pair(s, 60, parcels=4096)
    compare_state(g, o, 1e-4, "kerolox injector 160 with spray")
    PA, PB = g.spray.parcels()
    oA, oB = o.parcels(4096)
    alive = PB.reshape(-1, 4)[:, 1] > 0
    assert alive.sum() > 0
    assert np.array_equal(alive, oB.reshape(-1, 4)[:, 1] > 0)
    np.testing.assert_allclose(PA.reshape(-1, 4)[alive], oA.reshape(-1, 4)[alive], rtol=1e-3, atol=1e-3)
    st = g.spray.stats()
    assert st["parcels"] == int(alive.sum())


def test_injector_fp64_logic():
    s = SimSetup(configs.kerolox_injector(Nx=160))
    g, o = run_pair(s, 40, prec=64, parcels=4096)
    # parcels are stored as float32 on the device in both precisions
    compare_state(g, o, 1e-6, "kerolox injector FP64")


def test_rescale_ambient(cold):
    """rescaleAmbient mid-run, then 20 steps of the start-up transient, against the oracle
    at the fixed bar; the GPU FP32 result against the FP64 oracle is printed (the FP32
    oracle itself sits about 1e-4 from it in rho here: the transient amplifies rounding)."""
    g, o = run_pair(cold, 60)
    P2 = cold.compute_params(cold.cfg.Pc, cold.info.atmP)
    o64 = orc.Oracle(cold.Nx, cold.Nr, 64)
    A0, B0 = cold.initial_state()
    o64.set_geometry(cold.mask, cold.geoA, cold.geoG, cold.geoM, cold.jWall)
    o64.set_state(A0, B0)
    o64.step(60, cold.P)
    for x in (g, o, o64):
        (x.rescaleAmbient if x is g else x.rescale)(0.8, 0.95, 0.1, P2)
        x.step(20, P2)
    G = g.readAll()
    pg = prims(G["A"], G["B"])
    for k in pg:
        assert np.isfinite(pg[k]).all()
    gate(pg, prims(*o.state()), "after rescaleAmbient")
    gate(pg, prims(*o64.state_f64()), "after rescaleAmbient, GPU FP32 vs oracle FP64", bar=1e-2)


def test_readbacks_and_async(cold):
    g, o = run_pair(cold, 50)
    all_ = g.readAll()
    A = all_["A"].reshape(cold.Nr, cold.Nx, 4)
    col = g.readColumn(cold.iExit)
    np.testing.assert_array_equal(col["A"].reshape(cold.Nr, 4), A[:, cold.iExit])
    cell = g.readCell(cold.iChamber, 0)
    np.testing.assert_array_equal(cell["A"], A[0, cold.iChamber])
    # wall row (FS_WALLROW) and T2 (FS_VCELL) against the oracle
    g.drawWallRow()
    wr = g._readRect("wallRow", 0, 0, cold.Nx, 1)
    np.testing.assert_allclose(wr, o.wall_row(), rtol=1e-4, atol=1e-5)
    t2 = g._readRect("T2", 0, 0, cold.Nx, cold.Nr)
    ot2 = o.t2()
    assert rel_l2(t2.astype(np.float64), ot2.astype(np.float64)) < 1e-4
    reqs = [{"name": "step", "tex": "step", "x": 0, "y": 0, "w": 1, "h": 1},
            {"name": "colA", "tex": "A", "x": cold.iExit, "y": 0, "w": 1, "h": cold.Nr},
            {"name": "wall", "tex": "wallRow", "x": 0, "y": 0, "w": cold.Nx, "h": 1}]
    assert g.requestAsync(reqs, 7)
    assert not g.requestAsync(reqs, 8)
    res = None
    for _ in range(100000):
        res = g.pollAsync()
        if res:
            break
    assert res and res["tag"] == 7
    np.testing.assert_array_equal(res["data"]["colA"], col["A"])
    np.testing.assert_array_equal(res["data"]["step"], g.readStepInfo())


def test_integrals_match_host_sums(cold):
    g, _ = run_pair(cold, 300)
    P = cold.P
    I = g.integrals(P)
    col = g.readColumn(cold.iExit)
    A, B = col["A"].reshape(-1, 4).astype(np.float64), col["B"].reshape(-1, 4).astype(np.float64)
    dx = cold.dx
    pa = cold.probes().pa        # the ambient pressure at the double precision of the P block
    F = md = 0.0
    Ax = cold.Ax.reshape(cold.Nr + 1, cold.Nx + 1)
    for j in range(cold.jExit):
        rho = max(A[j, 0], 1e-9)
        u, v = A[j, 1] / rho, A[j, 2] / rho
        G = B[j, 2]
        p = max((A[j, 3] - 0.5 * rho * (u * u + v * v)) / G, 1e-9)
        dA = 2 * np.pi * abs(Ax[j, cold.iExit + 1]) * dx * dx
        F += (rho * u * u + p - pa) * dA
        md += B[j, 0] * u * dA
    assert I["F"] == pytest.approx(F * 101325.0, rel=1e-9, abs=1e-9)
    assert I["mdot"] == pytest.approx(md * 1.225 * np.sqrt(101325 / 1.225), rel=1e-6, abs=1e-12)
    assert I["mass"] > 0 and I["residual"] >= 0


@pytest.mark.parametrize("name", ["hot_bell", "cold_cone"])
def test_residual_matches_host(name):
    """The residual of a step (the RMS over fluid groups of the host's density change per
    unit grid time, over the chamber density; reduced on the device through per-tile
    partials folded across lanes and waves) against the same sum over the read-back states
    on the host: a fold that drops or doubles a lane, a row or a wave fails it."""
    s = SimSetup(getattr(configs, name)(Nx=384))
    g = GPUSolver()
    load_setup(g, s)
    g.step(60, s.P)
    a0 = g.readAll()["A"].reshape(-1, 4)[:, 0].copy()
    g.step(1, s.P)
    a1 = g.readAll()["A"].reshape(-1, 4)[:, 0].copy()
    dt = np.float32(g.readStepInfo()[0])
    res = g.integrals(s.P)["residual"]
    geoM = s.geoM.reshape(-1, 4)
    host = (s.mask.reshape(-1) == 0) & (geoM[:, 0] == 0) & (geoM[:, 1] == 0)
    dr = (a1[host] - a0[host]).astype(np.float64) / float(dt)
    scale = float(np.float32(float(s.P.res[2]) * float(s.P.res[1]) / float(s.P.res[0])))
    want = np.sqrt(np.sum(dr * dr) / host.sum()) * scale
    say(f"[parity] {name} 384: residual device {res:.9e} host {want:.9e} ({host.sum()} groups)")
    assert res > 0
    assert res == pytest.approx(want, rel=1e-9)
    g.close()


def plane_integrals(s, A, B):
    """Exit-plane thrust and exhaust mass flow (simsetup.js:314-337), throat mass flow
    (simsetup.js:239-250) and chamber pressure at (iChamber, 0) (app.js:537) of a state,
    in grid units, summed in float64 on the host."""
    A = A.reshape(s.Nr, s.Nx, 4).astype(np.float64)
    B = B.reshape(s.Nr, s.Nx, 4).astype(np.float64)
    Ax = s.Ax.reshape(s.Nr + 1, s.Nx + 1).astype(np.float64)
    mask = s.mask.reshape(s.Nr, s.Nx)

    def prim(j, i):
        rho = max(A[j, i, 0], 1e-9)
        u, v = A[j, i, 1] / rho, A[j, i, 2] / rho
        return rho, u, max((A[j, i, 3] - 0.5 * rho * (u * u + v * v)) / B[j, i, 2], 1e-9)

    F = md = mt = 0.0
    for j in range(s.jExit):
        rho, u, p = prim(j, s.iExit)
        dA = 2 * np.pi * abs(Ax[j, s.iExit + 1])
        F += (rho * u * u + p - s.P.amb[3]) * dA
        md += B[j, s.iExit, 0] * u * dA
    j = 0
    while j < s.Nr and not mask[j, s.iThroat]:
        rho, u, _ = prim(j, s.iThroat)
        mt += rho * u * 2 * np.pi * abs(Ax[j, s.iThroat])
        j += 1
    return {"F": F, "mdot": md, "mdotThroat": mt, "Pc": prim(0, s.iChamber)[2]}


def test_converged_integrals(cold):
    """Integrated thrust, exit and throat mass flow and chamber pressure on the converged
    cold-cone flow agree within 0.1 %:
    - FP32 against FP64 runs of the same kernels from the initial state, 15,000 steps
      (the flow is converged there: the FP64 integrals move by less than 0.1 % over the
      last 1,000 steps; the exit plane carries an unsteady over-expanded plume, about
      1e-4);
    - the HIP solver against the oracle, both started from the converged FP64 state
      (rounded to float32) and advanced 300 steps with the ignition ramp off (the state
      is past it; restarting the clock would otherwise replay the ramp), integrals and
      fields.
    Device integrals are the FP64 sums of the solver (nz_integrals_read); the oracle's
    state goes through the same host sums as the solver's."""
    keys = ("F", "mdot", "mdotThroat", "Pc")
    runs = {}
    for prec in (64, 32):
        g = GPUSolver(precision=prec)
        load_setup(g, cold)
        g.step(14000, cold.P)
        before = g.integrals(cold.P)
        g.step(1000, cold.P)
        runs[prec] = (g, before, g.integrals(cold.P))
    g64, b64, i64 = runs[64]
    _, _, i32 = runs[32]
    drift = {k: abs(i64[k] / b64[k] - 1) for k in keys}
    e32 = {k: abs(i32[k] / i64[k] - 1) for k in keys}
    say("[parity] cold cone 15000 steps: FP64 drift over the last 1000 steps " + " ".join(f"{k}={v:.1e}" for k, v in drift.items()))
    say("[parity] cold cone 15000 steps: FP32 vs FP64 integrals " + " ".join(f"{k}={v:.1e}" for k, v in e32.items()))
    assert max(drift.values()) < 1e-3
    assert max(e32.values()) < 1e-3

    # continuation from the converged state: HIP FP32 against the FP32 oracle
    st = g64.readAll()
    A, B = st["A"].astype(np.float32), st["B"].astype(np.float32)
    P = type(cold.P).from_buffer_copy(cold.P)
    P.ramp = 0.0
    g = GPUSolver()
    g.resize(cold.Nx, cold.Nr)
    g.setGeometry({"mask": cold.mask, "geoA": cold.geoA, "geoG": cold.geoG, "geoM": cold.geoM, "jWall": cold.jWall})
    g.setState(A, B)
    o = orc.Oracle(cold.Nx, cold.Nr, 32)
    o.set_geometry(cold.mask, cold.geoA, cold.geoG, cold.geoM, cold.jWall)
    o.set_state(A, B)
    g.step(300, P)
    o.step(300, P)
    G = g.readAll()
    Ao, Bo = o.state()
    ig, io = plane_integrals(cold, G["A"], G["B"]), plane_integrals(cold, Ao, Bo)
    eo = {k: abs(ig[k] / io[k] - 1) for k in keys}
    pg, po = prims(G["A"], G["B"]), prims(Ao, Bo)
    errs = {k: rel_l2(pg[k], po[k]) for k in pg}
    say("[parity] converged cold cone +300 steps: HIP vs oracle integrals " + " ".join(f"{k}={v:.1e}" for k, v in eo.items()))
    say("[parity] converged cold cone +300 steps: rel-L2 " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()))
    assert max(eo.values()) < 1e-3
    for k, v in errs.items():
        assert v <= 1e-4, (k, v)


@pytest.mark.timeout(400)
@pytest.mark.parametrize("name,t_end_ms,parcels", [("hot_bell", 2.0, None), ("kerolox_injector", 1.5, 65536)])
def test_horizon_integrals(name, t_end_ms, parcels):
    """The 768 x 256 workloads at their own horizons: the hot bell at 2 ms of simulated
    time (about 8250 steps), the kerolox injector with 65,536 parcels at 1.5 ms (about
    23,000 steps). The flows are unsteady by then, so FP32 and FP64 fields have drifted
    apart by a few percent in relative L2 (the FP32 oracle sits as far from the GPU), but
    the integrated quantities have not: exit-plane thrust and exhaust mass flow, throat
    mass flow and chamber pressure agree within 0.1 %, FP32 against FP64 runs of the same
    kernels and the GPU against the oracle over the first 8250 steps. The spray deposit
    sums in fixed point, so every run of the injector takes the same path."""
    s = SimSetup(getattr(configs, name)())
    keys = ("F", "mdot", "mdotThroat", "Pc")
    ORACLE_STEPS = 8250
    runs, snap = {}, None
    for prec in (32, 64):
        g = GPUSolver(precision=prec)
        if parcels:
            g.resize(s.Nx, s.Nr)
            g.spray.setCount(parcels)
        load_setup(g, s)
        steps = 0
        while float(g.readStepInfo()[2]) * s.tScale * 1e3 < t_end_ms:
            g.step(250, s.P)
            steps += 250
            if prec == 32 and steps == ORACLE_STEPS:
                snap = g.readAll()
        runs[prec] = (steps, g, g.integrals(s.P))
        if prec == 32 and snap is None:
            snap = g.readAll()
    steps, g32, i32 = runs[32]
    _, _, i64 = runs[64]
    e64 = {k: abs(i32[k] / i64[k] - 1) for k in keys}
    say(f"[parity] {name} {t_end_ms} ms ({steps} steps): FP32 vs FP64 integrals " + " ".join(f"{k}={v:.1e}" for k, v in e64.items()))
    assert max(e64.values()) < 1e-3
    o = orc.Oracle(s.Nx, s.Nr, 32)
    o.set_geometry(s.mask, s.geoA, s.geoG, s.geoM, s.jWall)
    if parcels:
        o.set_parcels(parcels)
    o.set_injector(s.injector)
    o.set_state(*s.initial_state())
    osteps = min(steps, ORACLE_STEPS)
    o.step(osteps, s.P)
    Ao, Bo = o.state()
    ig, io = plane_integrals(s, snap["A"], snap["B"]), plane_integrals(s, Ao, Bo)
    eo = {k: abs(ig[k] / io[k] - 1) for k in keys}
    say(f"[parity] {name} after {osteps} steps: GPU vs oracle integrals " + " ".join(f"{k}={v:.1e}" for k, v in eo.items()))
    assert max(eo.values()) < 1e-3


def test_deterministic_rerun(cold):
    g1, _ = run_pair(cold, 100)
    g2, _ = run_pair(cold, 100)
    a, b = g1.readAll(), g2.readAll()
    assert np.array_equal(a["A"], b["A"]) and np.array_equal(a["B"], b["B"])


@pytest.mark.parametrize("prec", [32, 64])
def test_spray_deterministic_rerun(prec):
    """The injector with 65,536 parcels twice from the same inputs, 300 steps: the spray
    deposit's sums are fixed-point integers, so the atomics' order leaves no trace and the
    gas state and every parcel come out bit for bit the same."""
    s = SimSetup(configs.kerolox_injector())
    out = []
    for _ in range(2):
        g = GPUSolver(precision=prec)
        g.resize(s.Nx, s.Nr)
        g.spray.setCount(65536)
        load_setup(g, s)
        g.step(300, s.P)
        st = g.readAllF64() if prec == 64 else g.readAll()
        out.append((st["A"], st["B"], *g.spray.parcels()))
    alive = int(np.sum(out[0][3].reshape(-1, 4)[:, 1] > 0))
    say(f"[parity] injector FP{prec} run twice, 300 steps, {alive} live parcels")
    assert alive > 1000
    for x, y in zip(*out):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("law", ["reference", "lognormal"])
def test_spawn_draws_on_device(law):
    """The parcels the device spawns in its first two steps against the draws of the
    literal pcg_hash table (tests/test_oracle_pins.py): element slot, angle, radius and
    diameter of each; with the reference's drop-size law and with the lognormal option
    (sigma 0.3 at 0.1 x the orifice)."""
    from test_oracle_pins import M32, lognormal_diameter, spray_rnd
    over = {"dropD0Frac": 0.1, "dropSigma": 0.3} if law == "lognormal" else {}
    s = SimSetup(configs.kerolox_injector(Nx=160, **over))
    g = GPUSolver()
    g.resize(s.Nx, s.Nr)
    g.spray.setCount(4096)
    load_setup(g, s)
    tab = s.injector.reshape(3, 256, 4).astype(np.float64)
    for step in range(2):
        g.step(1, s.P)
        dt = float(g.readStepInfo()[0])
        PA, PB = (x.reshape(-1, 4).astype(np.float64) for x in g.spray.parcels())
        seed = 1 + step
        for k in range(4):
            idx = 4 * step + k
            pid = (idx ^ ((seed * 2654435761) & M32)) & M32
            t = int(np.float32(np.float32(spray_rnd(pid, 1)) * np.float32(255.999)))
            ang = tab[0, t, 2] + (spray_rnd(pid, 2) - 0.5) * 2 * tab[1, t, 2]
            r = max(tab[0, t, 1] + (spray_rnd(pid, 3) - 0.5) * tab[1, t, 3], 0.05)
            if tab[2, t, 1] > 0:
                d = lognormal_diameter(tab[1, t, 1], tab[2, t, 1], pid)
            else:
                d = tab[1, t, 1] * np.exp((spray_rnd(pid, 4) - 0.5) * 1.4)
            a, b = PA[idx], PB[idx]
            assert a[0] == tab[0, t, 0] and b[2] == tab[1, t, 0] and b[3] == tab[2, t, 0], (idx, t)
            assert a[1] == pytest.approx(r, rel=1e-6)
            assert a[2] == pytest.approx(tab[0, t, 3] * np.cos(ang), rel=1e-5, abs=1e-6)
            assert a[3] == pytest.approx(tab[0, t, 3] * np.sin(ang), rel=1e-5, abs=1e-6)
            assert b[0] == pytest.approx(d, rel=1e-5)
            assert b[1] == pytest.approx(float(np.float32(s.P.sprayMhat)) * dt / 4, rel=1e-5)
    say("[parity] spawn draws on the device match the literal pcg_hash table")


def test_errors_are_raised():
    g = GPUSolver()
    with pytest.raises(Exception):
        g.step(1, SimSetup(configs.cold_cone(Nx=64)).P)   # no resize / geometry
    g.resize(64, 24)
    with pytest.raises(Exception):
        g._readRect("A", 60, 0, 10, 1)                   # outside the texture


U_REF = (287.053 * 288.15) ** 0.5   # sqrt(P_REF / RHO_REF), physics.js:11-15
DERIVED_FIELDS = {0: (250, 3600), 1: (0, 4), 2: (0, 1), 3: (3.5, 7), 4: (-2, 1.5), 5: (0, 3000), 6: (0, 2e5),
                  7: (0, 1), 9: (0, 3), 10: (0, 1), 11: (0, 1), 12: (0, 1), 13: (0, 6)}


@pytest.mark.parametrize("prec", [32, 64])
def test_derived_display_fields(prec):
    """FS_DERIVED on the device against the oracle's restatement, every field, after a
    short hot-bell run with afterburning (FP32 rel-L2 <= 1e-4 per channel; FP64 1e-9, and
    1e-7 for fields 9 and 13, which read vis.T2: its texels are float32 on both sides, and a
    texel whose FP64 value rounds to the neighbouring float moves by 6e-8)."""
    from rnozzle.solver import derived_params
    s = SimSetup(configs.hot_bell(Nx=192))
    g, o = run_pair(s, 60, prec=prec)
    tol = 1e-4 if prec == 32 else 1e-9
    for f, rng in DERIVED_FIELDS.items():
        out = g.derived(f, rng, s.P, 3500.0, s.info.atmT, 10.0, s.dx).reshape(-1, 4).astype(np.float64)
        ref = o.derived(derived_params(f, rng, s.P, 3500.0, s.info.atmT, 10.0, s.dx), U_REF).reshape(-1, 4)
        wall = ref[:, 3] < 0
        assert np.array_equal(out[:, 3] < 0, wall)
        for ch in range(4):
            e = rel_l2(out[~wall, ch], ref[~wall, ch].astype(np.float64))
            assert e <= (1e-7 if prec == 64 and f in (9, 13) else tol), (f, ch, e)
    say(f"[parity] derived fields FP{prec}: all {len(DERIVED_FIELDS)} fields within {tol}")


@pytest.mark.parametrize("prec,tol", [(32, None), (64, 1e-9)])
def test_live_remesh(prec, tol):
    """remesh() (app.js:254-270): a new geometry at the same cell size (a steeper 18-degree
    cone) handed to setGeometry on a running flow, which carries on. The GPU and the oracle
    take the same sequence: 150 steps, setGeometry, 60 steps; FP32 at the fixed bar, FP64 to
    1e-9."""
    from rnozzle import configs as cf
    s = SimSetup(cf.cold_cone(Nx=128))
    g, o = run_pair(s, 150, prec=prec)
    s2 = SimSetup(cf.apply(cf.cold_cone(Nx=128), {"coneAngle": 18.0, "dx": s.dx}))
    assert (s2.Nx, s2.Nr) == (s.Nx, s.Nr)
    assert not np.array_equal(s2.mask, s.mask)        # the wall moved
    g.setGeometry({"mask": s2.mask, "geoA": s2.geoA, "geoG": s2.geoG, "geoM": s2.geoM, "jWall": s2.jWall})
    o.set_geometry(s2.mask, s2.geoA, s2.geoG, s2.geoM, s2.jWall)
    g.step(60, s2.P)
    o.step(60, s2.P)
    g.sync()
    compare_state(g, o, tol if tol else 1e-4, f"cold cone re-meshed to 18 deg, FP{prec}")


@pytest.mark.parametrize("viscous", [0, 1])
def test_nan_heals_to_ambient(viscous):
    """A NaN planted in a running flow (setState) is healed in one step by the update's NaN
    guard (solver.js:348): the field is finite again and the GPU agrees with the oracle. On
    the inviscid flow every cell the guard resets holds exactly the ambient state (with the
    viscous split step on, the split step then acts on the reset cells as on any other)."""
    s = SimSetup(configs.apply(configs.cold_cone(Nx=128), {"viscous": viscous}))
    g, _ = run_pair(s, 100)
    st = g.readAll()
    A, B = st["A"].reshape(s.Nr, s.Nx, 4).copy(), st["B"].reshape(s.Nr, s.Nx, 4).copy()
    fluid = s.mask.reshape(s.Nr, s.Nx) == 0
    jn, i_n = 5, 40
    assert fluid[jn, i_n]
    A[jn, i_n, :] = np.nan
    g.setState(A.ravel(), B.ravel())
    o = orc.Oracle(s.Nx, s.Nr, 32)
    o.set_geometry(s.mask, s.geoA, s.geoG, s.geoM, s.jWall)
    o.set_state(A.ravel(), B.ravel())
    g.step(1, s.P)
    o.step(1, s.P)
    G = g.readAll()
    assert np.isfinite(G["A"]).all() and np.isfinite(G["B"]).all()
    Ao, Bo = o.state()
    assert np.isfinite(Ao).all() and np.isfinite(Bo).all()
    amb = np.array(list(s.P.ambCons), np.float32)
    healed = np.all(G["A"].reshape(s.Nr, s.Nx, 4) == amb, axis=-1) & fluid
    if not viscous:
        assert healed[jn, i_n], "the NaN cell did not take the ambient state"
        assert np.array_equal(healed, np.all(Ao.reshape(s.Nr, s.Nx, 4) == amb, axis=-1) & fluid)
    gate(prims(G["A"], G["B"]), prims(Ao, Bo), f"NaN planted at ({i_n}, {jn}), viscous={viscous}: "
         f"{int(healed.sum())} cells at the ambient state")
