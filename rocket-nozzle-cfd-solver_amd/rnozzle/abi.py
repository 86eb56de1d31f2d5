"""ctypes mirror of include/nozzle.h (the C ABI of the HIP solver library).

Structures here must match the header field for field; `check_layout()` in the
tests compares their sizes with what the library reports.
"""
import ctypes as C

f4 = C.c_float * 4
f3 = C.c_float * 3
f2 = C.c_float * 2


class NzParams(C.Structure):
    """The reference's normalised `P` block (simsetup.js:257-311) as float uniforms."""
    _fields_ = [
        ("amb", f4), ("res", f3), ("inlet", f4), ("pStart", C.c_float), ("ramp", C.c_float),
        ("RcCells", C.c_float), ("inj", C.c_int32),
        ("order2", C.c_int32), ("hllBlend", C.c_float),
        ("ambCons", f4), ("ambB", f4), ("gamR", f2), ("uInlet0", C.c_float), ("vMax", C.c_float),
        ("cfl", C.c_float), ("dtMax", C.c_float), ("wFloor", C.c_float),
        ("visc", C.c_int32), ("chem", C.c_int32), ("gas", f4), ("Rs", f3), ("chamberNu", f2),
        ("cap", f2), ("chemK", f4), ("Tw", C.c_float),
        ("sprayMhat", C.c_float), ("sprayTScale", C.c_float), ("fuelVec", f4), ("oxVec", f4),
        ("eF", C.c_float), ("eO", C.c_float), ("Ge", C.c_float),
    ]

    def copy(self):
        out = NzParams()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(NzParams))
        return out


# field order of nz_setup_get_params_f64 (and of the golden P vectors)
PARAM_ORDER = [
    ("amb", 4), ("res", 3), ("inlet", 4), ("pStart", 1), ("ramp", 1), ("RcCells", 1), ("inj", 1),
    ("order2", 1), ("hllBlend", 1), ("ambCons", 4), ("ambB", 4), ("gamR", 2), ("uInlet0", 1), ("vMax", 1),
    ("cfl", 1), ("dtMax", 1), ("wFloor", 1), ("visc", 1), ("chem", 1), ("gas", 4), ("Rs", 3),
    ("chamberNu", 2), ("cap", 2), ("chemK", 4), ("Tw", 1), ("sprayMhat", 1), ("sprayTScale", 1),
    ("fuelVec", 4), ("oxVec", 4), ("eF", 1), ("eO", 1), ("Ge", 1),
]
N_PARAMS_F64 = sum(n for _, n in PARAM_ORDER)


class NzConfig(C.Structure):
    """cfg of cfgFromS (app.js:201-211), SI units, plus build-defined extensions."""
    _fields_ = [
        ("Dt", C.c_double), ("De", C.c_double), ("type", C.c_int32),
        ("bellPct", C.c_double), ("coneAngle", C.c_double), ("CR", C.c_double), ("chamberLen", C.c_double),
        ("Pc", C.c_double), ("Tc", C.c_double), ("gamma", C.c_double), ("Mw", C.c_double),
        ("alt", C.c_double), ("mach", C.c_double), ("base", C.c_int32), ("bodyRatio", C.c_double),
        ("domainL", C.c_double), ("Nx", C.c_int32), ("aspect", C.c_double),
        ("startMode", C.c_int32), ("rampMs", C.c_double),
        ("cfl", C.c_double), ("order2", C.c_int32), ("wallModel", C.c_int32), ("minCells", C.c_int32),
        ("viscous", C.c_int32), ("afterburn", C.c_int32), ("qAb", C.c_double), ("sAir", C.c_double),
        ("Tw", C.c_double), ("inletAbsorb", C.c_double),
        ("prop", C.c_int32), ("injMode", C.c_int32), ("injType", C.c_int32), ("injRings", C.c_int32),
        ("injDP", C.c_double), ("injD", C.c_double), ("injTMR", C.c_double), ("injAngle", C.c_double), ("OF", C.c_double),
        ("film", C.c_double), ("baffles", C.c_double), ("baffleH", C.c_double), ("cavD", C.c_double), ("cavW", C.c_double),
        ("dx", C.c_double), ("hllBlend", C.c_double),
        ("convAngle", C.c_double), ("ambP", C.c_double), ("ambT", C.c_double), ("turbulence", C.c_int32),
        ("deviceGeometry", C.c_int32), ("dropD0Frac", C.c_double), ("dropSigma", C.c_double),
    ]


class NzCutcellIn(C.Structure):
    _fields_ = [
        ("Nx", C.c_int32), ("Nr", C.c_int32), ("dx", C.c_double), ("tw", C.c_double), ("xBase", C.c_double),
        ("bodyR", C.c_double), ("base", C.c_int32), ("stair", C.c_int32), ("Rtab", C.POINTER(C.c_float)),
        ("npts", C.c_int32), ("xs", C.POINTER(C.c_double)), ("rs", C.POINTER(C.c_double)),
        ("Rc", C.c_double), ("Re", C.c_double), ("xe", C.c_double), ("cavD", C.c_double), ("cavW", C.c_double),
        ("nsolids", C.c_int32), ("solids", C.POINTER(C.c_double)),
    ]


class NzSetupInfo(C.Structure):
    _fields_ = [
        ("Nx", C.c_int32), ("Nr", C.c_int32), ("dx", C.c_double), ("tScale", C.c_double), ("tw", C.c_double),
        ("bodyR", C.c_double), ("xBase", C.c_double),
        ("iExit", C.c_int32), ("jExit", C.c_int32), ("iChamber", C.c_int32), ("iThroat", C.c_int32),
        ("injector", C.c_int32), ("nCut", C.c_int32), ("nMerged", C.c_int32),
        ("Rt", C.c_double), ("Re", C.c_double), ("Rc", C.c_double), ("Lc", C.c_double), ("xt", C.c_double),
        ("xe", C.c_double), ("eps", C.c_double), ("thn", C.c_double), ("the", C.c_double),
        ("atmT", C.c_double), ("atmP", C.c_double),
        ("thMe", C.c_double), ("thPe", C.c_double), ("thMdot", C.c_double), ("thF", C.c_double),
        ("thIsp", C.c_double), ("thCstar", C.c_double), ("thC0", C.c_double), ("thUlim", C.c_double),
    ]


class NzReadReq(C.Structure):
    _fields_ = [("field", C.c_int32), ("x", C.c_int32), ("y", C.c_int32), ("w", C.c_int32), ("h", C.c_int32)]


class NzSprayStats(C.Structure):
    _fields_ = [("fuel", C.c_double), ("ox", C.c_double), ("parcels", C.c_int64), ("meanD", C.c_double)]


class NzProbes(C.Structure):
    _fields_ = [("iExit", C.c_int32), ("jExit", C.c_int32), ("iThroat", C.c_int32), ("iChamber", C.c_int32),
                ("dx", C.c_double), ("exact", C.c_int32), ("pa", C.c_double), ("Rs", C.c_double * 3),
                ("RaN", C.c_double), ("sprayMdot", C.c_double)]


class NzIntegrals(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("F", "mdot", "Me", "pe", "Isp", "mdotThroat", "Pc", "Tc", "mass", "energy", "residual",
                 "unburntOx", "fuelVapour")]


class NzDerivedParams(C.Structure):
    _fields_ = [("field", C.c_int32), ("range", C.c_float * 2), ("RaN", C.c_float), ("Rs", C.c_float * 3),
                ("chemK", C.c_float * 4), ("Tc", C.c_float), ("Ta", C.c_float), ("schlGain", C.c_float),
                ("dx", C.c_float)]


class NzSeparation(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "frac", "eps", "pRatio", "Msep")]


class NzBatchJob(C.Structure):
    _fields_ = [("Nx", C.c_int32), ("Nr", C.c_int32), ("mask", C.POINTER(C.c_uint8)),
                ("geoA", C.POINTER(C.c_float)), ("geoG", C.POINTER(C.c_float)), ("geoM", C.POINTER(C.c_float)),
                ("jWall", C.POINTER(C.c_float)), ("A", C.POINTER(C.c_float)), ("B", C.POINTER(C.c_float)),
                ("P", NzParams), ("injector", C.POINTER(C.c_float)), ("parcels", C.c_int32),
                ("probes", NzProbes), ("t_end", C.c_double), ("res_tol", C.c_double), ("min_steps", C.c_int32),
                ("max_steps", C.c_int32), ("keep_state", C.c_int32)]


class NzBatchResult(C.Structure):
    _fields_ = [("state", C.c_int32), ("stop", C.c_int32), ("steps", C.c_int64), ("t_grid", C.c_double),
                ("dt", C.c_double), ("residual", C.c_double), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("have_integrals", C.c_int32), ("integ", NzIntegrals)]


NZ_JOB_QUEUED, NZ_JOB_RUNNING, NZ_JOB_DONE = range(3)
NZ_STOP_NONE, NZ_STOP_RESIDUAL, NZ_STOP_TIME, NZ_STOP_STEPS = range(4)

NZ_FIELD_A, NZ_FIELD_B, NZ_FIELD_STEP, NZ_FIELD_WALLROW, NZ_FIELD_T2 = range(5)
NZ_BELL, NZ_CONE = 0, 1
NZ_START_RAMP, NZ_START_BURST = 0, 1
NZ_WALL_CUT, NZ_WALL_STAIR = 0, 1
PROPS = ["kerolox", "methalox", "hydrolox", "hybrid", "solid", "hydrazine", "coldgas", "custom"]
INJ_TYPES = ["shower", "coax", "impinge", "pintle"]

P = C.POINTER
vp = C.c_void_p
fp = P(C.c_float)
dp = P(C.c_double)
u8p = P(C.c_uint8)

# name -> (restype, argtypes); every function declared in include/nozzle.h
SIGNATURES = {
    "nz_last_error": (C.c_char_p, []),
    "nz_device_available": (C.c_int, []),
    "nz_create": (C.c_int, [P(vp)]),
    "nz_create_ex": (C.c_int, [P(vp), C.c_int]),
    "nz_destroy": (C.c_int, [vp]),
    "nz_resize": (C.c_int, [vp, C.c_int, C.c_int]),
    "nz_get_size": (C.c_int, [vp, P(C.c_int), P(C.c_int)]),
    "nz_set_mask": (C.c_int, [vp, u8p]),
    "nz_set_geometry": (C.c_int, [vp, u8p, fp, fp, fp, fp]),
    "nz_set_state": (C.c_int, [vp, fp, fp]),
    "nz_step": (C.c_int, [vp, C.c_int, P(NzParams)]),
    "nz_step_timed": (C.c_int, [vp, C.c_int, P(NzParams), dp]),
    "nz_rescale_ambient": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, P(NzParams)]),
    "nz_draw_wall_row": (C.c_int, [vp]),
    "nz_tile_shapes": (C.c_int, [C.c_int, P(C.c_int), P(C.c_int), P(C.c_int), C.c_int]),
    "nz_set_tile_shape": (C.c_int, [vp, C.c_int]),
    "nz_get_tile_shape": (C.c_int, [vp, P(C.c_int)]),
    "nz_autotune": (C.c_int, [vp, P(NzParams), C.c_int, P(C.c_int), dp]),
    "nz_read_rect": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, fp]),
    "nz_read_step_info": (C.c_int, [vp, fp]),
    "nz_read_state_f64": (C.c_int, [vp, dp, dp]),
    "nz_step_count": (C.c_int, [vp, P(C.c_int64)]),
    "nz_sync": (C.c_int, [vp]),
    "nz_read_async_begin": (C.c_int, [vp, C.c_int, P(NzReadReq), C.c_int32]),
    "nz_read_async_poll": (C.c_int, [vp, fp, P(C.c_int32)]),
    "nz_read_async_cancel": (C.c_int, [vp]),
    "nz_read_async_pending": (C.c_int, [vp]),
    "nz_spray_set_injector": (C.c_int, [vp, fp]),
    "nz_spray_reset": (C.c_int, [vp]),
    "nz_spray_stats": (C.c_int, [vp, P(NzSprayStats)]),
    "nz_spray_set_count": (C.c_int, [vp, C.c_int]),
    "nz_spray_read": (C.c_int, [vp, fp, fp]),
    "nz_set_probes": (C.c_int, [vp, P(NzProbes)]),
    "nz_integrals_read": (C.c_int, [vp, P(NzParams), P(NzIntegrals)]),
    "nz_read_derived": (C.c_int, [vp, P(NzDerivedParams), fp]),
    "nz_batch_create": (C.c_int, [P(vp), C.c_int, C.c_uint64, C.c_int]),
    "nz_batch_destroy": (C.c_int, [vp]),
    "nz_batch_instance_bytes": (C.c_uint64, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "nz_batch_submit": (C.c_int, [vp, P(NzBatchJob), P(C.c_int64)]),
    "nz_batch_run": (C.c_int, [vp, C.c_int, P(C.c_int), P(C.c_int)]),
    "nz_batch_gpu_busy": (C.c_int, [vp, P(C.c_double)]),
    "nz_batch_result_get": (C.c_int, [vp, C.c_int64, P(NzBatchResult)]),
    "nz_batch_read_state": (C.c_int, [vp, C.c_int64, fp, fp]),
    "nz_batch_release": (C.c_int, [vp, C.c_int64]),
    "nz_batch_set_graph": (C.c_int, [vp, C.c_int]),
    "nz_batch_memory": (C.c_int, [vp, P(C.c_uint64), P(C.c_uint64)]),
    "nz_batch_set_tile_shape": (C.c_int, [vp, C.c_int]),
    "nz_batch_autotune": (C.c_int, [vp, C.c_int, P(C.c_int), dp]),
    "nz_batch_finished": (C.c_int, [vp, P(C.c_int64), C.c_int]),
    "nz_config_default": (None, [P(NzConfig)]),
    "nz_config_preset": (C.c_int, [P(NzConfig), C.c_char_p]),
    "nz_setup_build": (C.c_int, [P(NzConfig), P(vp)]),
    "nz_setup_free": (None, [vp]),
    "nz_setup_get_info": (C.c_int, [vp, P(NzSetupInfo)]),
    "nz_setup_get_params": (C.c_int, [vp, P(NzParams)]),
    "nz_setup_get_params_f64": (C.c_int, [vp, dp, C.c_int]),
    "nz_setup_compute_params": (C.c_int, [vp, P(NzConfig), C.c_double, C.c_double, P(NzParams)]),
    "nz_setup_get_geometry": (C.c_int, [vp, u8p, fp, fp, fp, fp, fp, fp, fp]),
    "nz_setup_initial_state": (C.c_int, [vp, fp, fp]),
    "nz_setup_get_injector": (C.c_int, [vp, fp]),
    "nz_setup_get_probes": (C.c_int, [vp, P(NzProbes)]),
    "nz_setup_get_cutcell_input": (C.c_int, [vp, P(NzCutcellIn)]),
    "nz_cutcell_build": (C.c_int, [P(NzCutcellIn), u8p, fp, fp, fp, fp, fp, fp, P(C.c_int32)]),
    "nz_cutcell_build_noalloc": (C.c_int, [P(NzCutcellIn), u8p, fp, fp, fp, fp, fp, fp, P(C.c_int32)]),
    "nz_setup_radius_at": (C.c_double, [vp, C.c_double]),
    "nz_wall_separation": (C.c_int, [vp, fp, C.c_double, P(NzSeparation)]),
    "nz_summerfield_point": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, dp]),
    "nz_nozzle_theory": (C.c_int, [C.c_double] * 7 + [dp]),
    "nz_atmosphere": (None, [C.c_double, dp]),
}
