"""ctypes wrapper of the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py for the CPU-baseline timing. The product never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build_out", "liboracle.so")
REPO = os.path.dirname(HERE)

_lib = None


def build():
    """Compile oracle.cpp with g++ (OpenMP, no fast-math, no FMA contraction)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
           "-I", os.path.join(REPO, "include"), os.path.join(HERE, "oracle.cpp"), "-o", LIB + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: {LIB}")
        L = C.CDLL(LIB)
        vp, fp, dp = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double)
        sig = {
            "oracle_create": (vp, [C.c_int, C.c_int, C.c_int]),
            "oracle_destroy": (None, [vp]),
            "oracle_threads": (C.c_int, [C.c_int]),
            "oracle_set_geometry": (None, [vp, C.POINTER(C.c_uint8), fp, fp, fp, fp]),
            "oracle_set_state": (None, [vp, fp, fp]),
            "oracle_set_injector": (None, [vp, fp]),
            "oracle_set_parcels": (None, [vp, C.c_int]),
            "oracle_step": (None, [vp, C.c_int, C.c_void_p]),
            "oracle_rescale": (None, [vp, C.c_float, C.c_float, C.c_float, C.c_void_p]),
            "oracle_read_state": (None, [vp, fp, fp]),
            "oracle_read_state_f64": (None, [vp, dp, dp]),
            "oracle_read_step_info": (None, [vp, fp]),
            "oracle_read_t2": (None, [vp, fp]),
            "oracle_wall_row": (None, [vp, fp]),
            "oracle_read_parcels": (None, [vp, fp, fp]),
            "oracle_read_parcels_f64": (None, [vp, dp, dp]),
            "oracle_set_parcel": (None, [vp, C.c_int, fp, fp]),
            "oracle_derived": (None, [vp, C.c_void_p, C.c_double, fp]),
            "oracle_pcg": (C.c_int, [C.c_uint32]),
            "oracle_rnd": (C.c_float, [C.c_uint32, C.c_uint32]),
            "oracle_stage1_probe": (None, [vp, C.c_void_p, dp, dp, dp, dp, dp]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Oracle:
    """CPU restatement of GPUSolver + GPUSpray at FP32 (prec=32) or FP64 (prec=64)."""

    def __init__(self, Nx, Nr, prec=32, threads=0):
        self.Nx, self.Nr, self.prec = Nx, Nr, prec
        if threads:
            lib().oracle_threads(threads)
        self._h = lib().oracle_create(Nx, Nr, prec)

    def set_geometry(self, mask, geoA, geoG, geoM, jWall=None):
        m = np.ascontiguousarray(mask, np.uint8)
        arrs = [np.ascontiguousarray(a, np.float32) for a in (geoA, geoG, geoM)]
        jw = None if jWall is None else _f(np.ascontiguousarray(jWall, np.float32))
        self._keep = (m, arrs)
        lib().oracle_set_geometry(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8)), _f(arrs[0]), _f(arrs[1]), _f(arrs[2]), jw)

    def set_state(self, A, B):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        lib().oracle_set_state(self._h, _f(A), _f(B))

    def set_injector(self, tab):
        if tab is None:
            lib().oracle_set_injector(self._h, None)
        else:
            t = np.ascontiguousarray(tab, np.float32)
            lib().oracle_set_injector(self._h, _f(t))

    def set_parcels(self, n):
        lib().oracle_set_parcels(self._h, n)

    def step(self, n, P):
        lib().oracle_step(self._h, n, C.addressof(P))

    def rescale(self, fp, fT, du, P):
        lib().oracle_rescale(self._h, fp, fT, du, C.addressof(P))

    def state(self):
        A = np.zeros(self.Nx * self.Nr * 4, np.float32)
        B = np.zeros_like(A)
        lib().oracle_read_state(self._h, _f(A), _f(B))
        return A, B

    def state_f64(self):
        A = np.zeros(self.Nx * self.Nr * 4, np.float64)
        B = np.zeros_like(A)
        dp = C.POINTER(C.c_double)
        lib().oracle_read_state_f64(self._h, A.ctypes.data_as(dp), B.ctypes.data_as(dp))
        return A, B

    def step_info(self):
        o = np.zeros(4, np.float32)
        lib().oracle_read_step_info(self._h, _f(o))
        return o

    def t2(self):
        o = np.zeros(self.Nx * self.Nr * 4, np.float32)
        lib().oracle_read_t2(self._h, _f(o))
        return o

    def wall_row(self):
        o = np.zeros(self.Nx * 4, np.float32)
        lib().oracle_wall_row(self._h, _f(o))
        return o

    def set_parcel(self, k, a, b):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        lib().oracle_set_parcel(self._h, k, _f(a), _f(b))

    def derived(self, d, vel_scale):
        """FS_DERIVED fields of the current state; d is an nz_derived_params structure."""
        out = np.zeros(self.Nx * self.Nr * 4, np.float32)
        lib().oracle_derived(self._h, C.addressof(d), vel_scale, _f(out))
        return out

    def stage1_probe(self, P):
        """Stage 1 of the next step without committing it: (A1, B1) as (Nr, Nx, 4), the x-
        and r-face flux textures as (Nr + 1, Nx + 1, 12) = (F, FY, FO), and the step info."""
        Nx, Nr = self.Nx, self.Nr
        A1 = np.zeros(Nx * Nr * 4)
        B1 = np.zeros_like(A1)
        Fx = np.zeros((Nx + 1) * (Nr + 1) * 12)
        Fr = np.zeros_like(Fx)
        info = np.zeros(4)
        dp = C.POINTER(C.c_double)
        lib().oracle_stage1_probe(self._h, C.addressof(P), *(a.ctypes.data_as(dp) for a in (A1, B1, Fx, Fr, info)))
        return (A1.reshape(Nr, Nx, 4), B1.reshape(Nr, Nx, 4), Fx.reshape(Nr + 1, Nx + 1, 12),
                Fr.reshape(Nr + 1, Nx + 1, 12), info)

    def parcels(self, n):
        PA = np.zeros(n * 4, np.float32)
        PB = np.zeros(n * 4, np.float32)
        lib().oracle_read_parcels(self._h, _f(PA), _f(PB))
        return PA, PB

    def parcels_f64(self, n):
        PA = np.zeros(n * 4)
        PB = np.zeros(n * 4)
        dp = C.POINTER(C.c_double)
        lib().oracle_read_parcels_f64(self._h, PA.ctypes.data_as(dp), PB.ctypes.data_as(dp))
        return PA, PB

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().oracle_destroy(self._h)
                self._h = None
        except Exception:
            pass
