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


@pytest.mark.parametrize("name,precision", [("hot_bell", 32), ("cold_cone", 32), ("hot_bell", 64)])
def test_every_shape_bit_identical(name, precision):
    s = SimSetup(configs.hot_bell(Nx=192) if name == "hot_bell" else configs.cold_cone())
    shapes = tile_shapes(precision)
    assert len(shapes) >= 3 and shapes[0] in ((64, 4), (32, 4))
    ref = run(s, 40, 0, precision)
    for k in range(1, len(shapes)):
        say(f"[shapes] {name} FP{precision} shape {k} {shapes[k]}")
        out = run(s, 40, k, precision)
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
