"""Python view of the native host setup (buildSimSetup and friends, simsetup.js:7-311).

`SimSetup(cfg)` builds one solver instance's inputs with the C++ restatement in
csrc/setup.cpp and exposes them as numpy arrays in the reference's texture
layouts (row-major, RGBA-interleaved float32).
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import check, lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def default_config(**over):
    """cfg of cfgFromS with the app's default settings (app.js:12-26, 201-211)."""
    c = abi.NzConfig()
    lib().nz_config_default(C.byref(c))
    return apply_config(c, **over)


def preset_config(name, **over):
    """cfg after applyPreset(name) (app.js:1108-1121)."""
    c = default_config()
    check(lib().nz_config_preset(C.byref(c), name.encode()), "nz_config_preset")
    return apply_config(c, **over)


def apply_config(c, **over):
    for k, v in over.items():
        if k == "type" and isinstance(v, str):
            v = {"bell": abi.NZ_BELL, "cone": abi.NZ_CONE}[v]
        elif k == "prop" and isinstance(v, str):
            v = abi.PROPS.index(v)
        elif k == "injType" and isinstance(v, str):
            v = abi.INJ_TYPES.index(v)
        elif k == "startMode" and isinstance(v, str):
            v = {"ramp": abi.NZ_START_RAMP, "burst": abi.NZ_START_BURST}[v]
        elif k == "wallModel" and isinstance(v, str):
            v = {"cut": abi.NZ_WALL_CUT, "stair": abi.NZ_WALL_STAIR}[v]
        elif isinstance(v, bool):
            v = int(v)
        if not hasattr(c, k):
            raise KeyError(f"unknown config field {k}")
        setattr(c, k, v)
    return c


class SimSetup:
    """One instance's grid, geometry, parameters and initial state."""

    def __init__(self, cfg):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().nz_setup_build(C.byref(cfg), C.byref(h)), "nz_setup_build")
        self._h = h
        info = abi.NzSetupInfo()
        check(lib().nz_setup_get_info(h, C.byref(info)), "nz_setup_get_info")
        self.info = info
        self.Nx, self.Nr, self.dx = info.Nx, info.Nr, info.dx
        Nx, Nr = self.Nx, self.Nr
        self.mask = np.zeros(Nx * Nr, np.uint8)
        self.geoA = np.zeros((Nx + 1) * (Nr + 1) * 4, np.float32)
        self.geoG = np.zeros(Nx * Nr * 4, np.float32)
        self.geoM = np.zeros(Nx * Nr * 4, np.float32)
        self.jWall = np.zeros(Nx * 4, np.float32)
        self.kappa = np.zeros(Nx * Nr, np.float32)
        self.Ax = np.zeros((Nx + 1) * (Nr + 1), np.float32)
        self.Ar = np.zeros((Nx + 1) * (Nr + 1), np.float32)
        check(lib().nz_setup_get_geometry(
            h, self.mask.ctypes.data_as(C.POINTER(C.c_uint8)), _fptr(self.geoA), _fptr(self.geoG),
            _fptr(self.geoM), _fptr(self.jWall), _fptr(self.kappa), _fptr(self.Ax), _fptr(self.Ar)),
            "nz_setup_get_geometry")
        self.P = abi.NzParams()
        check(lib().nz_setup_get_params(h, C.byref(self.P)), "nz_setup_get_params")
        self.P64 = np.zeros(abi.N_PARAMS_F64, np.float64)
        lib().nz_setup_get_params_f64(h, self.P64.ctypes.data_as(C.POINTER(C.c_double)), abi.N_PARAMS_F64)
        self.injector = None
        if info.injector:
            tab = np.zeros(256 * 3 * 4, np.float32)
            check(lib().nz_setup_get_injector(h, _fptr(tab)), "nz_setup_get_injector")
            self.injector = tab
        self.iExit, self.jExit, self.iChamber, self.iThroat = info.iExit, info.jExit, info.iChamber, info.iThroat
        self.tScale = info.tScale

    def initial_state(self):
        A = np.zeros(self.Nx * self.Nr * 4, np.float32)
        B = np.zeros(self.Nx * self.Nr * 4, np.float32)
        check(lib().nz_setup_initial_state(self._h, _fptr(A), _fptr(B)), "nz_setup_initial_state")
        return A, B

    def compute_params(self, PcSi, pStartSi, cfg=None):
        P = abi.NzParams()
        check(lib().nz_setup_compute_params(self._h, C.byref(cfg or self.cfg), PcSi, pStartSi, C.byref(P)),
              "nz_setup_compute_params")
        return P

    def radius_at(self, x):
        return lib().nz_setup_radius_at(self._h, x)

    def probes(self):
        """Probes with the parameter block's double-precision values (nz_setup_get_probes)."""
        pr = abi.NzProbes()
        check(lib().nz_setup_get_probes(self._h, C.byref(pr)), "nz_setup_get_probes")
        return pr

    def wall_separation(self, row, paSi):
        """wallSeparation(setup, row, paSi) (simsetup.js:352-373) on the near-wall row
        (Nx x (u, v, p, c), as nz_draw_wall_row leaves it): a dict, or None while attached."""
        row = np.ascontiguousarray(row, np.float32)
        if row.size < self.Nx * 4:
            raise ValueError(f"wall row needs {self.Nx * 4} floats, got {row.size}")
        out = abi.NzSeparation()
        if lib().nz_wall_separation(self._h, _fptr(row), float(paSi), C.byref(out)) != 1:
            return None
        return {n: getattr(out, n) for n, _ in abi.NzSeparation._fields_}

    def summerfield_point(self, PcSi, gamma, paSi):
        """summerfieldPoint(contour, Pc, g, pa) (simsetup.js:375-384): fraction of the bell,
        or None if the 1-D wall pressure never drops below 0.4 pa."""
        frac = C.c_double()
        if lib().nz_summerfield_point(self._h, float(PcSi), float(gamma), float(paSi), C.byref(frac)) != 1:
            return None
        return frac.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().nz_setup_free(self._h)
                self._h = None
        except Exception:
            pass
