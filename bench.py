"""Solver-step throughput on one MI355X: cell-updates/s and HBM roofline fraction.

A step is one full solver time step over the whole grid (device CFL time step,
SSP-RK2 stages 1 and 2, the viscous / wall / afterburning split step, and the
spray when the workload has one), on inputs already resident in HBM.

Default workload: hot methalox bell, 768 x 256, FP32, full viscous model
(`rnozzle.configs.hot_bell`). Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--workload hot_bell|cold_cone|kerolox_injector]
  python bench.py --workload batch [--instances 64]   (batched solves/s vs one at a time)
  python bench.py --workload serve [--requests 64]    (Poisson arrivals, time-to-converged)
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
# FP64 runs of the same kernels hold the state in doubles: the 8-plane state streams double
# (stage 1: 2 x 32 B more, stage 2: 3 x 32 B more), geometry and sources stay float32
STATE_STREAMS = {1: 2, 2: 3}


def stage_bytes(stage, precision):
    base = STAGE1_BYTES if stage == 1 else STAGE2_BYTES
    return base + (STATE_STREAMS[stage] * 32 if precision == 64 else 0)
PARCEL_BYTES = 64 + 32 + 16   # parcel read+write, gas gather, 4-float deposit

# HBM-side bytes per launch of the stage-2 kernel measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE (scaled by a 256 MiB copy: x2.0 and x1.0), for (workload, grid, tile shape);
# profiles/hot_bell_768x256_counters.md holds the runs
MEASURED_TRAFFIC = {("hot_bell", "768x256", 1): 27.86e6}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(setup, budget_s=15.0, budget1_s=5.0, precision=32):
    """The oracle (FP32 CPU restatement, OpenMP) on the same workload for a bounded
    number of steps; cell-updates/s on this host's cores, and on one core beside it."""
    import oracle
    # the host's CPU share: OMP_NUM_THREADS where the environment sets it, else every
    # core this process may run on (nproc)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0)) or 1

    def timed(nthreads, budget, chunk):
        o = oracle.Oracle(setup.Nx, setup.Nr, precision, threads=nthreads)
        o.set_geometry(setup.mask, setup.geoA, setup.geoG, setup.geoM, setup.jWall)
        o.set_injector(setup.injector)
        A, B = setup.initial_state()
        o.set_state(A, B)
        o.step(2, setup.P)  # warm the pages
        steps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < budget:
            o.step(chunk, setup.P)
            steps += chunk
        return steps, time.perf_counter() - t0

    steps, dt = timed(threads, budget_s, 5)
    steps1, dt1 = timed(1, budget1_s, 1)
    cells = setup.Nx * setup.Nr
    return {"value": cells * steps / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP{precision} on {setup.Nx}x{setup.Nr}, {steps} steps in {dt:.1f} s ({cpu_model()}, "
                      f"nproc {os.cpu_count()})",
            "one_core": {"value": cells * steps1 / dt1, "sample": f"{steps1} steps in {dt1:.1f} s on 1 thread"}}


def pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v, float), q)) if len(v) else None


def plain_one_at_a_time(items):
    """The same solves through GPUSolver, one after another, each stepped exactly the
    step count its batched run took (nz_step in chunks, one synchronisation at the end),
    so the plain path does no stop-test polling at all."""
    from rnozzle.solver import GPUSolver, load_setup
    total, cu = 0.0, 0.0
    for n, (s, steps) in enumerate(items):
        if n % 8 == 0:
            print(f"[bench] plain solve {n}/{len(items)}", file=sys.stderr, flush=True)
        g = GPUSolver()
        if s.injector is not None:
            g.resize(s.Nx, s.Nr)
            g.spray.setCount(16384)
        load_setup(g, s)
        g.sync()
        t0 = time.perf_counter()
        left = steps
        while left > 0:
            n = min(left, 1000)
            g.step(n, s.P)
            left -= n
        g.sync()
        total += time.perf_counter() - t0
        cu += s.Nx * s.Nr * steps
        g.close()
    return total, cu


def tune_shape(setup, args):
    """Tile shape for a batch: the autotuner's pick on its largest solve (0 with --no-tune)."""
    if args.shape >= 0:
        return args.shape
    if args.no_tune:
        return 0
    from rnozzle.solver import GPUSolver, load_setup
    g = GPUSolver()
    load_setup(g, setup)
    g.step(50, setup.P)
    best, _ = g.autotune(setup.P, steps=20)
    g.close()
    return best


def bench_batch(args):
    """Batched solve set: `--instances` nozzles drawn from `--seed`, each to a residual
    of 1e-6 or 3 ms simulated, all resident and advancing in the same launches;
    against the same solves one at a time through GPUSolver."""
    import numpy as np
    from rnozzle import configs
    from rnozzle.batch import Batch
    from rnozzle.sim_setup import SimSetup
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    setups = [SimSetup(configs.batch_instance(rng)) for _ in range(args.instances)]
    t_setup = time.perf_counter() - t0
    b = Batch(max_resident=max(1, args.instances), budget_bytes=int(args.budget_gb * (1 << 30)))
    ids = [b.submit(s, t_end_s=configs.BATCH_T_END, res_tol=configs.BATCH_RES_TOL, min_steps=100) for s in setups]
    t0 = time.perf_counter()
    # the autotuner runs inside the timed region: its trial steps advance the solves
    if args.shape >= 0:
        b.set_tile_shape(args.shape)
        shape, tune_us = args.shape, None
    else:
        shape, tune_us = b.autotune(20) if not args.no_tune else (0, None)
    runs = 0
    while True:
        r, q = b.run(args.chunk)
        runs += 1
        if runs % 20 == 0:
            print(f"[batch] {time.perf_counter() - t0:.1f} s: {r} resident, {q} queued", file=sys.stderr, flush=True)
        if r == 0 and q == 0:
            break
    wall = time.perf_counter() - t0
    res = [b.result(j) for j in ids]
    steps = [r["steps"] for r in res]
    cu = sum(s.Nx * s.Nr * n for s, n in zip(setups, steps))
    lat = [r["t_done"] - r["t_submit"] for r in res]
    stops = {"residual": sum(r["stop"] == 1 for r in res), "time": sum(r["stop"] == 2 for r in res)}
    out = {
        "metric": "batched solves/s", "value": len(ids) / wall, "unit": "solves/s", "n_gpus": 1,
        "steps": int(sum(steps)), "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": f"synthetic: {len(ids)} nozzles drawn from seed {args.seed}",
        "config": {"workload": f"batch{len(ids)}", "instances": len(ids), "seed": args.seed,
                   "t_end_s": configs.BATCH_T_END, "res_tol": configs.BATCH_RES_TOL, "chunk_steps": args.chunk,
                   "tile_shape": shape, "autotune_us_per_step": tune_us,
                   "grids": sorted({f"{s.Nx}x{s.Nr}" for s in setups})},
        "wall_s": wall, "cell_updates_per_s": cu / wall, "setup_build_s": t_setup, "stops": stops,
        "latency_s": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": max(lat)},
        "gpu_busy_fraction": b.gpu_busy() / wall,
    }
    if args.plain != 0:
        n = len(ids) if args.plain < 0 else min(args.plain, len(ids))
        tp, cup = plain_one_at_a_time(list(zip(setups[:n], steps[:n])))
        out["plain"] = {"solves": n, "wall_s": tp, "solves_per_s": n / tp, "cell_updates_per_s": cup / tp,
                        "how": "GPUSolver per solve, same step counts, sequential"}
        out["speedup_vs_plain"] = out["value"] / (n / tp)
    return out


def bench_serve(args):
    """Serving: Poisson arrivals (rate `--rate`/s) of `--requests` requests, 70 %
    nozzles as in the batched set, 30 % 384x128 spray injectors with 16,384 parcels.
    Arrived requests wait on the host; at most `--max-active` are in the runner at once
    (admitted there by memory budget), picked cheapest-first by estimated cell-steps
    (sjf) or in arrival order (fifo). Latency = completion - arrival."""
    import numpy as np
    from rnozzle import configs
    from rnozzle.batch import Batch
    from rnozzle.sim_setup import SimSetup
    rng = np.random.default_rng(args.seed)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, args.requests))
    reqs = []
    for k in range(args.requests):
        inj = rng.uniform() < 0.3
        cfg = configs.injector_instance(rng) if inj else configs.batch_instance(rng)
        reqs.append((float(arrivals[k]), SimSetup(cfg), inj))
    shape = tune_shape(max((r[1] for r in reqs), key=lambda s: s.Nx * s.Nr), args)
    b = Batch(max_resident=args.requests, budget_bytes=int(args.budget_gb * (1 << 30)))
    b.set_tile_shape(shape)
    def work(s):
        # estimated cell-steps to the simulated horizon: cells x horizon / (CFL dt at the wave-speed floor)
        return s.Nx * s.Nr * configs.BATCH_T_END / s.tScale * float(s.P.wFloor) / float(s.P.cfl)

    t0 = time.perf_counter()
    m0 = time.monotonic()
    arrival_of = {}
    nxt = 0
    waiting = []   # arrived, not yet handed to the runner
    done = []
    while True:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt][0] <= now:
            waiting.append(reqs[nxt])
            nxt += 1
        # admission: keep at most --max-active solves in the runner, cheapest first (sjf) or in order
        if args.policy == "sjf":
            waiting.sort(key=lambda r: work(r[1]))
        r, q = b.run(0)
        while waiting and r + q < args.max_active:
            a, s, inj = waiting.pop(0)
            jid = b.submit(s, t_end_s=configs.BATCH_T_END, res_tol=configs.BATCH_RES_TOL, min_steps=100,
                           parcels=configs.INJECTOR_PARCELS if inj else None)
            arrival_of[jid] = (a, inj)
            q += 1
        r, q = b.run(args.chunk)
        done += b.finished()
        if nxt == len(reqs) and not waiting and r == 0 and q == 0:
            break
        if r == 0 and q == 0 and not waiting and nxt < len(reqs):
            time.sleep(max(0.0, reqs[nxt][0] - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    # the runner's clock is the host steady clock (CLOCK_MONOTONIC), as time.monotonic()
    lat, lat_inj, lat_noz = [], [], []
    for jid in done:
        res = b.result(jid)
        a, inj = arrival_of[jid]
        l = res["t_done"] - (m0 + a)
        lat.append(l)
        (lat_inj if inj else lat_noz).append(l)
    out = {
        "metric": "serving solves/s", "value": len(done) / wall, "unit": "solves/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": f"synthetic: {args.requests} Poisson arrivals at {args.rate}/s from seed {args.seed}",
        "config": {"workload": f"serve{args.requests}", "rate_per_s": args.rate, "requests": args.requests,
                   "injector_share": 0.3, "parcels": configs.INJECTOR_PARCELS, "t_end_s": configs.BATCH_T_END,
                   "res_tol": configs.BATCH_RES_TOL, "chunk_steps": args.chunk, "policy": args.policy,
                   "max_active": args.max_active, "tile_shape": shape},
        "wall_s": wall,
        "time_to_converged_s": {"p50": pct(lat, 50), "p99": pct(lat, 99),
                                "nozzle_p50": pct(lat_noz, 50), "injector_p50": pct(lat_inj, 50)},
        # device time of the run chunks over the wall time from the first arrival to the last result
        "gpu_busy_fraction": b.gpu_busy() / wall,
    }
    return out


def calibrate_copy(mib):
    """A plain device copy of known size (torch's copy kernel), run 3 times: the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters of a profile."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": n * 4, "bytes_written": n * 4, "launches": 3}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="hot_bell")
    ap.add_argument("--precision", type=int, default=32, choices=[32, 64],
                    help="arithmetic of the run (64: the FP64 error-reference run of the same kernels)")
    ap.add_argument("--nx", type=int, default=0, help="override the grid width (other sizes are parity cases)")
    ap.add_argument("--set", action="append", default=[], help="config override key=value (e.g. wallModel=stair)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shape")
    ap.add_argument("--shape", type=int, default=-1, help="fix the tile shape id (no autotuning); for profiles")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--develop-ms", type=float, default=1.0,
                    help="simulated ms run untimed before the warm-up, so the timed steps see a developed flow")
    ap.add_argument("--no-batch", action="store_true", help="headline only: skip the batch and serve blocks")
    ap.add_argument("--instances", type=int, default=64, help="batch: solves in the set")
    ap.add_argument("--seed", type=int, default=20240501)
    ap.add_argument("--chunk", type=int, default=200, help="batch/serve: steps per run call")
    ap.add_argument("--plain", type=int, default=-1, help="batch: solves re-run one at a time (-1 all, 0 none)")
    ap.add_argument("--budget-gb", type=float, default=16.0)
    ap.add_argument("--rate", type=float, default=200.0, help="serve: mean arrivals per second")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--policy", default="sjf", choices=["sjf", "fifo"], help="serve: admission order")
    ap.add_argument("--max-active", type=int, default=16, help="serve: solves in the runner at once")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run only a device copy of MIB MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported (independent solves "
              "would shard one queue per GPU; a single grid would split along x with a 2-column halo)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return

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
    g = GPUSolver(precision=args.precision)
    load_setup(g, s)
    cells = s.Nx * s.Nr
    spray = s.injector is not None
    # develop the flow (untimed): the timed steps then run on the plume, not the start-up
    def develop(t_ms, at_least=0):
        n = 0
        while n < at_least or float(g.readStepInfo()[2]) * s.tScale * 1e3 < t_ms:
            g.step(500 if n >= at_least else min(500, at_least - n), s.P)
            n += 500 if n >= at_least else min(500, at_least - n)
        return n
    develop_steps = develop(0.9 * args.develop_ms)
    # tile shape: the autotuner times every shape from the current state (restored after
    # each trial) and keeps the fastest; outside the timed region
    from rnozzle.solver import shape_name, tile_shapes
    tune_us = None
    if args.shape >= 0:
        g.setTileShape(args.shape)
    elif not args.no_tune:
        _, tune_us = g.autotune(s.P, steps=20)
    shape = g.tileShape()
    # the rest of the development at the chosen shape, at least 400 steps of it: the GPU
    # slows its clocks while it waits on the host between the autotuner's trials, and takes
    # several milliseconds of continuous work to come back to full speed
    if args.develop_ms > 0:
        develop_steps += develop(args.develop_ms, at_least=400)
    # W warm-up steps at the chosen shape, right before the timed region
    g.step(args.warmup, s.P)
    g.sync()
    t_start_ms = float(g.readStepInfo()[2]) * s.tScale * 1e3
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
    bpc2 = stage_bytes(2, args.precision) + (STAGE2_SPRAY_EXTRA if spray else 0)
    bpc1 = stage_bytes(1, args.precision)
    b2, b1 = cells * bpc2, cells * bpc1
    tx, ty = tile_shapes(args.precision)[shape]
    rtype = "float" if args.precision == 32 else "double"
    roof = {"bound": "hbm", "kernel": f"k_stage<{rtype},{tx},{ty},2,{'true' if s.P.visc or s.P.chem else 'false'}> "
                                      "(flux + update + split step + CFL partials, fused)",
            "achieved": b2 / t2 / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b2 / t2 / 1e9 / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get((args.workload, f"{s.Nx}x{s.Nr}", shape))
            if not over and args.precision == 32 else None,
            "bytes_per_cell": bpc2,
            "avg_us": t2 * 1e6,
            "stage1": {"achieved": b1 / t1 / 1e9, "frac": b1 / t1 / 1e9 / HBM_PEAK_GBS, "avg_us": t1 * 1e6,
                       "bytes_per_cell": bpc1}}
    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if args.precision == 32 else "fp64",
        "data": "synthetic (setup from the workload's geometry)",
        "config": {"workload": args.workload, "overrides": over, "grid": f"{s.Nx}x{s.Nr}", "cells": cells,
                   "merged_cells": int(s.info.nMerged), "physics": bool(s.P.visc),
                   "afterburning": bool(s.P.chem), "spray": spray, "cfl": round(float(s.P.cfl), 3),
                   "tile": shape_name(shape, args.precision),
                   "autotune_us_per_step": None if tune_us is None else
                   {shape_name(k, args.precision): round(u, 2) for k, u in enumerate(tune_us)}},
        "roofline": roof,
        "kernel_ms": {"steps": ktimed, "stage1": ms[0], "spray": ms[1], "stage2": ms[2], "events_span": ms[3]},
        "sim": {"t_start_ms": t_start_ms, "develop_steps": develop_steps, "t_ms": float(si[2]) * s.tScale * 1e3,
                "dt": float(si[0]), "Pc_bar": integ["Pc"] / 1e5, "mdot_exit": integ["mdot"], "thrust_N": integ["F"],
                "mdot_throat": integ["mdotThroat"], "residual": integ["residual"]},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(s, args.cpu_seconds, precision=args.precision)
    g.close()
    if not args.no_batch and args.workload == "hot_bell" and args.precision == 32 and not args.nx and not over:
        # the batched and serving workloads beside the headline (their own timed regions)
        print("[bench] batch: 64 solves batched, then one at a time", file=sys.stderr, flush=True)
        bt = bench_batch(args)
        out["batch"] = {k: bt[k] for k in ("value", "unit", "wall_s", "cell_updates_per_s", "stops", "latency_s",
                                           "gpu_busy_fraction")}
        out["batch"]["config"] = bt["config"]
        if "plain" in bt:
            out["batch"]["plain"] = bt["plain"]
            out["batch"]["speedup_vs_plain"] = bt["speedup_vs_plain"]
        print("[bench] serve: Poisson arrivals", file=sys.stderr, flush=True)
        sv = bench_serve(args)
        out["serve"] = {k: sv[k] for k in ("value", "unit", "wall_s", "time_to_converged_s", "gpu_busy_fraction")}
        out["serve"]["config"] = sv["config"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
