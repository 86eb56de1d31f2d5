"""Tile shapes of the stage kernels and the autotuner.

Every face flux and every group update is evaluated by the same arithmetic whatever
the tiling (faces of merged cells outside a tile included), so every tile shape must
give bit-identical states; the autotuner must leave the solve where it found it.
"""
import sys

import numpy as np
import pytest

from rnozzle import configs
from rnozzle._lib import lib
from rnozzle.sim_setup import SimSetup
from rnozzle.solver import GPUSolver, load_setup, tile_shapes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if not lib().nz_device_available():
        pytest.fail("no gfx950 device: the HIP path must run on an MI355X")


def run(setup, steps, shape, precision=32):
    g = GPUSolver(precision=precision)
    g.resize(setup.Nx, setup.Nr)
    g.setTileShape(shape)
    load_setup(g, setup)
    assert g.tileShape() == shape
    g.step(steps, setup.P)
    return g.readAllF64() if precision == 64 else g.readAll()


@pytest.mark.parametrize("name,precision", [("hot_bell", 32), ("cold_cone", 32), ("hot_bell", 64), ("hot_bell_200", 32),
                                            ("kerolox_injector", 32)])
def test_every_shape_bit_identical(name, precision):
    # hot_bell_200 (200 x 67) and the injector (160 x 53) leave partial tiles at the grid's
    # east and north edges for most shapes
    if name == "hot_bell":
        s = SimSetup(configs.hot_bell(Nx=192))
    elif name == "hot_bell_200":
        s = SimSetup(configs.hot_bell(Nx=200))
    elif name == "kerolox_injector":
        s = SimSetup(configs.kerolox_injector(Nx=160))
    else:
        s = SimSetup(configs.cold_cone())
    shapes = tile_shapes(precision)
    assert len(shapes) >= 3 and shapes[0] in ((64, 4), (32, 4))
    ref = run(s, 40, 0, precision)
    for k in range(1, len(shapes)):
        say(f"[shapes] {name} FP{precision} shape {k} {shapes[k]}")
        out = run(s, 40, k, precision)
        assert np.array_equal(out["A"], ref["A"]) and np.array_equal(out["B"], ref["B"]), shapes[k]


@pytest.mark.parametrize("order2,hll", [(0, 1.0), (1, 0.0)])
def test_every_shape_bit_identical_scheme_options(order2, hll):
    # first-order faces (the per-cell MUSCL values formed but not taken) and no HLL blend
    # (the jump ratios not formed): the per-cell and the stencil face forms still agree
    s = SimSetup(configs.hot_bell(Nx=192))
    s.P.order2 = order2
    s.P.hllBlend = hll
    shapes = tile_shapes(32)
    ref = run(s, 40, 0)
    for k in range(1, len(shapes)):
        say(f"[shapes] hot_bell order2={order2} hllBlend={hll} shape {k} {shapes[k]}")
        out = run(s, 40, k)
        assert np.array_equal(out["A"], ref["A"]) and np.array_equal(out["B"], ref["B"]), shapes[k]


def test_autotune_leaves_state_and_picks_a_shape():
    s = SimSetup(configs.hot_bell(Nx=384))
    g = GPUSolver()
    load_setup(g, s)
    g.step(20, s.P)
    best, us = g.autotune(s.P, steps=10)
    say(f"[shapes] autotune 384: best {best} {tile_shapes()[best]}, us/step " + " ".join(f"{u:.1f}" for u in us))
    assert g.tileShape() == best and all(u > 0 for u in us) and us[best] == min(us)
    g.step(20, s.P)
    a = g.readAll()
    h = GPUSolver()
    load_setup(h, s)
    h.step(40, s.P)
    b = h.readAll()
    assert np.array_equal(a["A"], b["A"]) and np.array_equal(a["B"], b["B"])
    assert g.stepCount == 40


def test_autotune_restores_spray_state():
    """On a spray solve the autotuner's trial steps move parcels and deposit sources; it
    restores fields, parcels and step info after each trial, so the solve reads back
    exactly as before the call."""
    s = SimSetup(configs.kerolox_injector(Nx=160))
    g = GPUSolver()
    g.resize(s.Nx, s.Nr)
    g.spray.setCount(4096)
    g.setGeometry({"mask": s.mask, "geoA": s.geoA, "geoG": s.geoG, "geoM": s.geoM, "jWall": s.jWall})
    g.spray.setInjector({"tab": s.injector})
    g.setState(*s.initial_state())
    g.step(20, s.P)
    before, pa0, si0 = g.readAll(), g.spray.parcels(), g.readStepInfo()
    g.autotune(s.P, steps=5)
    after, pa1, si1 = g.readAll(), g.spray.parcels(), g.readStepInfo()
    assert np.array_equal(before["A"], after["A"]) and np.array_equal(before["B"], after["B"])
    assert np.array_equal(pa0[0], pa1[0]) and np.array_equal(pa0[1], pa1[1])
    assert np.array_equal(si0, si1)


def test_timed_steps_match_plain_steps():
    """nz_step_timed (the benchmark's per-kernel timing run) launches the same kernels with
    events around each: the state it leaves is bit-identical to nz_step's."""
    import ctypes as C
    from rnozzle._lib import check
    s = SimSetup(configs.hot_bell(Nx=192))
    g1, g2 = GPUSolver(), GPUSolver()
    for g in (g1, g2):
        load_setup(g, s)
    g1.step(30, s.P)
    ms = (C.c_double * 4)()
    check(lib().nz_step_timed(g2._h, 30, C.byref(s.P), ms), "nz_step_timed")
    a, b = g1.readAll(), g2.readAll()
    assert np.array_equal(a["A"], b["A"]) and np.array_equal(a["B"], b["B"])
    assert ms[0] > 0 and ms[2] > 0 and ms[3] >= ms[0]
