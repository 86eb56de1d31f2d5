"""rnozzle — MI355X-native (gfx950 HIP) rocket-nozzle solver hot path.

Python host of the C ABI in include/nozzle.h: the solver (`GPUSolver`, mirror
of the reference's solver.js class), the host setup (`SimSetup`), and the
batched / serving runtime.
"""
from . import abi  # noqa: F401
from ._lib import NozzleError, lib  # noqa: F401
