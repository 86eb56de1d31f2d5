"""GPUSolver / GPUSpray over the HIP C ABI, with the reference's method names.

Mirrors the JavaScript classes of source/solver.js:677-1076 and
source/spray.js:211-321 (same methods, arguments, return shapes and error
behaviour: failures raise, `step` never waits for the device). Arrays are numpy
float32 in the reference's RGBA texture layout.
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import NozzleError, check, lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _req_floats(reqs):
    return sum(r["w"] * r["h"] * 4 for r in reqs)


class GPUSpray:
    """GPUSpray (spray.js:211-321): parcels live on the device with the solver."""

    def __init__(self, solver):
        self.s = solver
        self.blendOk = True      # fixed-point integer atomics replace EXT_float_blend (spray.js:215)
        self.ready = True
        self.n = 65536

    def resize(self):
        self.reset()

    def setInjector(self, inj):
        if inj is None:
            return
        tab = inj["tab"] if isinstance(inj, dict) else inj
        t = np.ascontiguousarray(tab, np.float32)
        if t.size != 256 * 3 * 4:
            raise NozzleError("setInjector: table must hold 256*3*4 floats")
        check(lib().nz_spray_set_injector(self.s._h, _f(t)), "setInjector")

    def reset(self):
        check(lib().nz_spray_reset(self.s._h), "spray.reset")

    def setCount(self, n):
        check(lib().nz_spray_set_count(self.s._h, int(n)), "spray.setCount")
        self.n = int(n)

    def stats(self, P=None):
        st = abi.NzSprayStats()
        check(lib().nz_spray_stats(self.s._h, C.byref(st)), "spray.stats")
        return {"fuel": st.fuel, "ox": st.ox, "parcels": int(st.parcels), "meanD": st.meanD}

    def parcels(self):
        PA = np.zeros(self.n * 4, np.float32)
        PB = np.zeros(self.n * 4, np.float32)
        check(lib().nz_spray_read(self.s._h, _f(PA), _f(PB)), "spray.read")
        return PA, PB


def derived_params(field, rng, P, Tc, Ta, schlGain, dx):
    """nz_derived_params as app.js fills the view (app.js:379-382): gas constants and
    kinetics from the parameter block P, chamber / ambient temperature in K."""
    d = abi.NzDerivedParams()
    d.field = int(field)
    d.range[:] = [float(rng[0]), float(rng[1])]
    d.RaN = P.gas[0]
    d.Rs[:] = list(P.Rs)
    d.chemK[:] = list(P.chemK)
    d.Tc, d.Ta, d.schlGain, d.dx = float(Tc), float(Ta), float(schlGain), float(dx)
    return d


def tile_shapes(precision=32, waves=False):
    """[(TX, TY)] (or [(TX, TY, W)] with waves=True) of the stage kernels' tile shapes at
    this precision; index = shape id."""
    n = lib().nz_tile_shapes(int(precision), None, None, None, 0)
    tx, ty, w = (C.c_int * n)(), (C.c_int * n)(), (C.c_int * n)()
    lib().nz_tile_shapes(int(precision), tx, ty, w, n)
    return list(zip(tx[:n], ty[:n], w[:n])) if waves else list(zip(tx[:n], ty[:n]))


def shape_name(shape, precision=32):
    tx, ty, w = tile_shapes(precision, waves=True)[shape]
    return f"{tx}x{ty}" + (f"w{w}" if w > 1 else "")


class GPUSolver:
    """GPUSolver (solver.js:677-1076) on one MI355X (HIP stream per solver)."""

    FIELDS = {"A": abi.NZ_FIELD_A, "B": abi.NZ_FIELD_B, "step": abi.NZ_FIELD_STEP,
              "wallRow": abi.NZ_FIELD_WALLROW, "T2": abi.NZ_FIELD_T2}

    def __init__(self, gl=None, precision=32):
        h = C.c_void_p()
        check(lib().nz_create_ex(C.byref(h), int(precision)), "GPUSolver")
        self._h = h
        self.precision = precision
        self.Nx = 0
        self.Nr = 0
        self.stepCount = 0
        self.spray = GPUSpray(self)
        self._async = None
        self.mask = None

    # ---- setup (solver.js:759-828)
    def resize(self, Nx, Nr):
        check(lib().nz_resize(self._h, int(Nx), int(Nr)), "resize")
        self.Nx, self.Nr = int(Nx), int(Nr)
        self._async = None

    def setMask(self, u8):
        m = np.ascontiguousarray(u8, np.uint8)
        check(lib().nz_set_mask(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8))), "setMask")
        self.mask = m

    def setGeometry(self, geo):
        m = np.ascontiguousarray(geo["mask"], np.uint8)
        gA = np.ascontiguousarray(geo["geoA"], np.float32)
        gG = np.ascontiguousarray(geo["geoG"], np.float32)
        gM = np.ascontiguousarray(geo["geoM"], np.float32)
        N, NL = self.Nx * self.Nr, (self.Nx + 1) * (self.Nr + 1)
        if m.size != N or gA.size != NL * 4 or gG.size != N * 4 or gM.size != N * 4:
            raise NozzleError("setGeometry: array sizes do not match the grid")
        jw = geo.get("jWall")
        jwp = None
        if jw is not None:
            jw = np.ascontiguousarray(jw, np.float32)
            jwp = _f(jw)
        check(lib().nz_set_geometry(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8)), _f(gA), _f(gG), _f(gM), jwp),
              "setGeometry")
        self.mask = m
        self.geoM = gM
        self.geoG = gG

    def setState(self, A, B):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        if A.size != self.Nx * self.Nr * 4 or B.size != A.size:
            raise NozzleError("setState: arrays must hold Nx*Nr*4 floats")
        check(lib().nz_set_state(self._h, _f(A), _f(B)), "setState")
        self.stepCount = 0
        self._async = None

    # ---- stepping (solver.js:849-891, 998-1010)
    def step(self, n, P):
        check(lib().nz_step(self._h, int(n), C.byref(P)), "step")
        self.stepCount += int(n)

    def rescaleAmbient(self, fp, fT, du, P):
        check(lib().nz_rescale_ambient(self._h, float(fp), float(fT), float(du), C.byref(P)), "rescaleAmbient")

    # ---- tile shapes of the stage kernels and the autotuner (no reference counterpart)
    def setTileShape(self, shape):
        check(lib().nz_set_tile_shape(self._h, int(shape)), "setTileShape")

    def tileShape(self):
        v = C.c_int()
        check(lib().nz_get_tile_shape(self._h, C.byref(v)), "tileShape")
        return v.value

    def autotune(self, P, steps=20):
        """Time every tile shape from the current state (left unchanged) and keep the
        fastest; returns (best shape, [us per step of each shape])."""
        n = len(tile_shapes(self.precision))
        us = (C.c_double * n)()
        best = C.c_int()
        check(lib().nz_autotune(self._h, C.byref(P), int(steps), C.byref(best), us), "autotune")
        return best.value, list(us)

    def drawWallRow(self):
        check(lib().nz_draw_wall_row(self._h), "drawWallRow")
        return "wallRow"

    def sync(self):
        check(lib().nz_sync(self._h), "sync")

    # ---- read-back (solver.js:1012-1075)
    @property
    def state(self):
        return {"A": "A", "B": "B"}

    @property
    def stepInfoTex(self):
        return "step"

    def _readRect(self, tex, x, y, w, h):
        out = np.zeros(w * h * 4, np.float32)
        check(lib().nz_read_rect(self._h, self.FIELDS[tex], x, y, w, h, _f(out)), "_readRect")
        return out

    def readStepInfo(self):
        return self._readRect("step", 0, 0, 1, 1)

    def readColumn(self, i):
        return {"A": self._readRect("A", i, 0, 1, self.Nr), "B": self._readRect("B", i, 0, 1, self.Nr)}

    def readCell(self, i, j):
        return {"A": self._readRect("A", i, j, 1, 1), "B": self._readRect("B", i, j, 1, 1)}

    def readAll(self):
        return {"A": self._readRect("A", 0, 0, self.Nx, self.Nr), "B": self._readRect("B", 0, 0, self.Nx, self.Nr)}

    def readAllF64(self):
        """The state at the solver's own precision (FP64 runs), RGBA-interleaved doubles."""
        n = self.Nx * self.Nr * 4
        A, B = np.zeros(n), np.zeros(n)
        dp = C.POINTER(C.c_double)
        check(lib().nz_read_state_f64(self._h, A.ctypes.data_as(dp), B.ctypes.data_as(dp)), "readAllF64")
        return {"A": A, "B": B}

    def requestAsync(self, reqs, tag):
        """reqs: [{name, tex, x, y, w, h}]; False while a read is in flight (solver.js:1028-1048)."""
        if self._async is not None:
            return False
        arr = (abi.NzReadReq * len(reqs))()
        for k, r in enumerate(reqs):
            arr[k] = abi.NzReadReq(self.FIELDS[r["tex"]], r["x"], r["y"], r["w"], r["h"])
        rc = lib().nz_read_async_begin(self._h, len(reqs), arr, int(tag))
        if rc < 0:
            check(rc, "requestAsync")
        if rc == 0:
            return False
        self._async = {"reqs": reqs, "tag": tag}
        return True

    def pollAsync(self):
        a = self._async
        if a is None:
            return None
        out = np.zeros(_req_floats(a["reqs"]), np.float32)
        tag = C.c_int32()
        rc = lib().nz_read_async_poll(self._h, _f(out), C.byref(tag))
        if rc < 0:
            check(rc, "pollAsync")
        if rc == 0:
            return None
        self._async = None
        data, off = {}, 0
        for r in a["reqs"]:
            n = r["w"] * r["h"] * 4
            data[r["name"]] = out[off:off + n]
            off += n
        return {"tag": a["tag"], "data": data, "probe": a.get("probe")}

    def cancelAsync(self):
        lib().nz_read_async_cancel(self._h)
        self._async = None

    # ---- build-defined extras
    def setProbes(self, iExit, jExit=None, iThroat=None, iChamber=None, dx=None):
        """Probes of the integrated quantities: an NzProbes (SimSetup.probes(), with the
        parameter block's double-precision values), or the four indices and dx."""
        pr = iExit if isinstance(iExit, abi.NzProbes) else abi.NzProbes(iExit, jExit, iThroat, iChamber, dx)
        check(lib().nz_set_probes(self._h, C.byref(pr)), "setProbes")

    def derived(self, field, rng, P, Tc, Ta, schlGain, dx):
        """render.js derive() (FS_DERIVED, render.js:60-134): per cell (t, Z, Tn, M)."""
        d = derived_params(field, rng, P, Tc, Ta, schlGain, dx)
        out = np.zeros(self.Nx * self.Nr * 4, np.float32)
        check(lib().nz_read_derived(self._h, C.byref(d), _f(out)), "derived")
        return out

    def integrals(self, P):
        out = abi.NzIntegrals()
        check(lib().nz_integrals_read(self._h, C.byref(P), C.byref(out)), "integrals")
        return {n: getattr(out, n) for n, _ in abi.NzIntegrals._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().nz_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_setup(solver, setup):
    """The restart() sequence of app.js:229-252 for a SimSetup."""
    solver.resize(setup.Nx, setup.Nr)
    solver.setGeometry({"mask": setup.mask, "geoA": setup.geoA, "geoG": setup.geoG, "geoM": setup.geoM,
                        "jWall": setup.jWall})
    if setup.injector is not None:
        solver.spray.setInjector({"tab": setup.injector})
    A, B = setup.initial_state()
    solver.setState(A, B)
    solver.setProbes(setup.probes())
    return A, B
