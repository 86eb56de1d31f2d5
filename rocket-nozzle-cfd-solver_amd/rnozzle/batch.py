"""Batched solves with continuous admission over the nz_batch_* C ABI.

Many independent solves of different grid sizes share every kernel launch; jobs
are admitted by device-memory budget as earlier ones finish. The reference has no
counterpart (one GPUSolver per page, app.js:151-167); a solve's result is the
same as a GPUSolver's on the same inputs (bit-identical without spray).
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import check, lib


def _f(a):
    """float* of a float32 C-contiguous array (the caller keeps `a` alive)."""
    assert a is None or (a.dtype == np.float32 and a.flags.c_contiguous)
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class Batch:
    def __init__(self, precision=32, budget_bytes=16 << 30, max_resident=256, graph=True):
        """`budget_bytes`: device memory for all residents, allocated at once as one slab."""
        h = C.c_void_p()
        check(lib().nz_batch_create(C.byref(h), precision, int(budget_bytes), max_resident), "nz_batch_create")
        self._h = h
        self.precision = precision
        self._setups = {}
        if not graph:
            check(lib().nz_batch_set_graph(h, 0), "nz_batch_set_graph")

    @staticmethod
    def instance_bytes(Nx, Nr, precision=32, parcels=0):
        return int(lib().nz_batch_instance_bytes(Nx, Nr, precision, parcels))

    def submit(self, setup, t_end_s=None, res_tol=0.0, min_steps=0, max_steps=0, keep_state=False, parcels=None,
               P=None, state=None):
        """Queue one solve of a SimSetup; returns its id. `t_end_s` is simulated
        seconds (converted with the setup's time unit dx / U_REF)."""
        A, B = state if state is not None else setup.initial_state()
        arrays = [np.ascontiguousarray(setup.mask, np.uint8)]
        arrays += [np.ascontiguousarray(x, np.float32) for x in (setup.geoA, setup.geoG, setup.geoM, setup.jWall, A, B)]
        tab = None if setup.injector is None else np.ascontiguousarray(setup.injector, np.float32)
        j = abi.NzBatchJob()
        j.Nx, j.Nr = setup.Nx, setup.Nr
        j.mask = arrays[0].ctypes.data_as(C.POINTER(C.c_uint8))
        j.geoA, j.geoG, j.geoM, j.jWall, j.A, j.B = (_f(a) for a in arrays[1:])
        j.P = P if P is not None else setup.P
        j.injector = _f(tab)
        j.parcels = (parcels or 65536) if tab is not None else 0
        j.probes = setup.probes()
        j.t_end = (t_end_s / setup.tScale) if t_end_s else 0.0
        j.res_tol = res_tol or 0.0
        j.min_steps = min_steps
        j.max_steps = max_steps
        j.keep_state = 1 if keep_state else 0
        jid = C.c_int64()
        check(lib().nz_batch_submit(self._h, C.byref(j), C.byref(jid)), "nz_batch_submit")
        self._setups[jid.value] = (setup.Nx, setup.Nr)
        return jid.value

    def run(self, steps):
        """Admit, enqueue `steps` steps of every resident, then read the previous run's
        stop flags and retire what stopped (the device runs this call's steps meanwhile);
        returns (solves not yet filed, queued jobs). A loop of calls ends at (0, 0)."""
        r, q = C.c_int(), C.c_int()
        check(lib().nz_batch_run(self._h, steps, C.byref(r), C.byref(q)), "nz_batch_run")
        return r.value, q.value

    def gpu_busy(self):
        """Device seconds of the run chunks read so far (event-timed)."""
        v = C.c_double()
        check(lib().nz_batch_gpu_busy(self._h, C.byref(v)), "nz_batch_gpu_busy")
        return v.value

    def finished(self):
        n = lib().nz_batch_finished(self._h, None, 0)
        ids = (C.c_int64 * max(n, 1))()
        lib().nz_batch_finished(self._h, ids, n)
        return list(ids[:n])

    def result(self, jid):
        r = abi.NzBatchResult()
        check(lib().nz_batch_result_get(self._h, jid, C.byref(r)), "nz_batch_result_get")
        out = {k: getattr(r, k) for k, _ in abi.NzBatchResult._fields_ if k != "integ"}
        out["integ"] = {k: getattr(r.integ, k) for k, _ in abi.NzIntegrals._fields_} if r.have_integrals else None
        return out

    def read_state(self, jid):
        Nx, Nr = self._setups[jid]
        A = np.empty(Nx * Nr * 4, np.float32)
        B = np.empty(Nx * Nr * 4, np.float32)
        check(lib().nz_batch_read_state(self._h, jid, _f(A), _f(B)), "nz_batch_read_state")
        return A, B

    def release(self, jid):
        check(lib().nz_batch_release(self._h, jid), "nz_batch_release")
        self._setups.pop(jid, None)

    def set_tile_shape(self, shape):
        check(lib().nz_batch_set_tile_shape(self._h, int(shape)), "nz_batch_set_tile_shape")

    def autotune(self, steps=20):
        """Run `steps` steps of the live batch with every tile shape and keep the fastest;
        returns (best shape, [us per step of each shape])."""
        n = lib().nz_tile_shapes(self.precision, None, None, None, 0)
        us = (C.c_double * n)()
        best = C.c_int()
        check(lib().nz_batch_autotune(self._h, int(steps), C.byref(best), us), "nz_batch_autotune")
        return best.value, list(us)

    def memory(self):
        u, b = C.c_uint64(), C.c_uint64()
        check(lib().nz_batch_memory(self._h, C.byref(u), C.byref(b)), "nz_batch_memory")
        return u.value, b.value

    def close(self):
        if getattr(self, "_h", None):
            lib().nz_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
