"""Solver-step throughput on one MI355X: cell-updates/s and HBM roofline fraction.

A step is one full solver time step over the whole grid (device CFL time step,
SSP-RK2 stages 1 and 2, the viscous / wall / afterburning split step, and the
spray when the workload has one), on inputs already resident in HBM.

Default workload: hot methalox bell, 768 x 256, FP32, full viscous model
(`rnozzle.configs.hot_bell`). Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--workload hot_bell|cold_cone|kerolox_injector]
"""
import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "rocket-nozzle-cfd-solver_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))

HBM_PEAK_GBS = 8000.0

# Compulsory HBM bytes per cell per launch of each kernel (FP32, SoA):
#   stage 1: read U^n (8 x 4 B) + write U^(1) (32 B) + geometry (ax, ar, 4 B each; group
#            V, Am, Wx, Wr 16 B; code word 4 B) = 92 B
#   stage 2: read U^(1) 32 + read U^n 32 + write U^{n+1} 32 + geometry 28 + eddy length 4
#            (+ spray sources read and cleared 32 B when the spray runs) = 128 (160) B
STAGE1_BYTES = 92
STAGE2_BYTES = 128
STAGE2_SPRAY_EXTRA = 32
PARCEL_BYTES = 64 + 32 + 16   # parcel read+write, gas gather, 4-float deposit


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(setup, budget_s=15.0):
    """The oracle (FP32 CPU restatement, OpenMP) on the same workload for a bounded
    number of steps; cell-updates/s on this host's cores."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = min(threads, 16)
    o = oracle.Oracle(setup.Nx, setup.Nr, 32, threads=threads)
    o.set_geometry(setup.mask, setup.geoA, setup.geoG, setup.geoM, setup.jWall)
    o.set_injector(setup.injector)
    A, B = setup.initial_state()
    o.set_state(A, B)
    o.step(2, setup.P)  # warm the pages
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        o.step(5, setup.P)
        steps += 5
    dt = time.perf_counter() - t0
    return {"value": setup.Nx * setup.Nr * steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 on {setup.Nx}x{setup.Nr}, {steps} steps in {dt:.1f} s ({cpu_model()})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="hot_bell")
    ap.add_argument("--nx", type=int, default=0, help="override the grid width (other sizes are parity cases)")
    ap.add_argument("--set", action="append", default=[], help="config override key=value (e.g. wallModel=stair)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(instances shard across GPUs by design, see the sharding notes)", file=sys.stderr)
        sys.exit(2)

    from rnozzle import configs
    from rnozzle._lib import check, lib
    from rnozzle.sim_setup import SimSetup
    from rnozzle.solver import GPUSolver, load_setup

    make = getattr(configs, args.workload)
    cfg = make(Nx=args.nx) if args.nx else make()
    over = {}
    for kv in args.set:
        k, v = kv.split("=", 1)
        try:
            v = float(v) if "." in v or "e" in v.lower() else int(v)
        except ValueError:
            pass
        over[k] = v
    if over:
        cfg = configs.apply(cfg, over)
    s = SimSetup(cfg)
    g = GPUSolver()
    load_setup(g, s)
    cells = s.Nx * s.Nr
    spray = s.injector is not None
    g.step(args.warmup, s.P)
    g.sync()
    # timed region: K steps enqueued back to back on the solver's stream
    t0 = time.perf_counter()
    g.step(args.steps, s.P)
    g.sync()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    value = cells * args.steps / wall
    # per-kernel durations: the same K steps again with HIP events around every launch
    # on the solver's stream (the events add launch gaps, so `value` comes from the run above)
    ktimed = min(args.steps, 1000)
    ms = (C.c_double * 4)()
    check(lib().nz_step_timed(g._h, ktimed, C.byref(s.P), ms), "nz_step_timed")
    si = g.readStepInfo()
    integ = g.integrals(s.P)
    # dominant kernel: the fused stage-2 kernel (stage 1 + split step + spray apply + CFL)
    t2 = ms[2] / ktimed * 1e-3
    t1 = ms[0] / ktimed * 1e-3
    b2 = cells * (STAGE2_BYTES + (STAGE2_SPRAY_EXTRA if spray else 0))
    b1 = cells * STAGE1_BYTES
    roof = {"bound": "hbm", "kernel": "k_stage<2> (flux + update + split step + CFL, fused)",
            "achieved": b2 / t2 / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b2 / t2 / 1e9 / HBM_PEAK_GBS,
            "traffic": None, "bytes_per_cell": STAGE2_BYTES + (STAGE2_SPRAY_EXTRA if spray else 0),
            "avg_us": t2 * 1e6,
            "stage1": {"achieved": b1 / t1 / 1e9, "frac": b1 / t1 / 1e9 / HBM_PEAK_GBS, "avg_us": t1 * 1e6,
                       "bytes_per_cell": STAGE1_BYTES}}
    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (setup from the workload's geometry)",
        "config": {"workload": args.workload, "overrides": over, "grid": f"{s.Nx}x{s.Nr}", "cells": cells,
                   "merged_cells": int(s.info.nMerged), "physics": bool(s.P.visc),
                   "afterburning": bool(s.P.chem), "spray": spray, "cfl": round(float(s.P.cfl), 3)},
        "roofline": roof,
        "kernel_ms": {"steps": ktimed, "stage1": ms[0], "spray": ms[1], "stage2": ms[2], "events_span": ms[3]},
        "sim": {"t_ms": float(si[2]) * s.tScale * 1e3, "dt": float(si[0]), "Pc_bar": integ["Pc"] / 1e5,
                "mdot_exit": integ["mdot"], "thrust_N": integ["F"], "residual": integ["residual"]},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
