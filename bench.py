#!/usr/bin/env python3
"""Benchmark of the per-level Godunov hydro path on one MI355X.

Workload (default): single-level roofline bench -- 2^21 octs (256^3 =
16,777,216 leaf cells) at one level of the reference's uniform octree
(reference level 9, periodic through the tree), seeded state rho, p ~
lognormal(0, 0.5), v_i ~ N(0, 0.5 c_s), gamma = 5/3, HLLC, MinMod.  A step is
one Simulation::amr_step of that level: Courant reduction, set_unew,
godunov_fine, set_uold, all on the device with the state resident in HBM.

Prints one JSON line (metric of the benchmark spec, roofline of the godunov
kernel from HIP events on its stream, CPU baseline = the oracle timed on a
bounded sample of the same state on this host).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ramses-cpp_amd"))

METRIC = "hydro cell-updates/s on 1 MI355X, FP64 vs mixed+batched; % of HBM roofline"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 80  # read uold (5 x FP64) + write unew (5 x FP64)


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--level", type=int, default=9, help="reference level list (9 -> 256^3 cells)")
    ap.add_argument("--riemann", default="hllc")
    ap.add_argument("--slope", type=int, default=1)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (one GPU per run here)",
              file=sys.stderr)
        return 2

    import numpy as np

    import rhydro

    gamma = 5.0 / 3.0
    level = args.level
    t_setup = time.time()
    mesh = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(mesh, level, args.seed, gamma, 0.5, 0.5)
    print(f"[bench] mesh level {level}: {mesh.ngridmax} grids, {mesh.ncell} cells, setup "
          f"{time.time() - t_setup:.1f}s", file=sys.stderr, flush=True)

    s = rhydro.Solver(gamma=gamma, slope_type=args.slope, riemann=args.riemann, precision=rhydro.RH_PREC_FP64)
    s.bind(mesh, rhydro.RH_DEVICE_RESIDENT)
    nleaf = s.leaf_cells(level)
    dx = 1.0 / (2 ** level)  # Simulation.cpp:548, boxlen/(nx*2^ilevel)
    cf = 0.8

    s.run_uniform(level, dx, cf, args.warmup)  # ends with a stream synchronisation
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    s.timers(True)
    s.timers_reset()
    t0 = time.perf_counter()
    dts = s.run_uniform(level, dx, cf, args.steps)  # synchronises before returning
    t1 = time.perf_counter()
    wall = t1 - t0
    g_ms, g_n = s.timer(rhydro.TK_GODUNOV)
    c_ms, _ = s.timer(rhydro.TK_COURANT)
    u_ms, _ = s.timer(rhydro.TK_SET_UNEW)
    o_ms, _ = s.timer(rhydro.TK_SET_UOLD)
    s.timers(False)
    print(f"[bench] timed {args.steps} steps: {wall:.3f}s; dt[0..2]={dts[:3].tolist()}", file=sys.stderr,
          flush=True)

    value = nleaf * args.steps / wall
    god_avg_s = (g_ms / max(g_n, 1)) / 1e3
    achieved = BYTES_PER_CELL_UPDATE * nleaf / god_avg_s / 1e9

    cpu = None
    if not args.no_cpu and args.cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        s.sync_to_host()
        s.close()
        om = oracle.OracleMesh(mesh)
        p = oracle.params(gamma, args.slope, rhydro.resolve_riemann(args.riemann), 1, 0, 1.0)
        sample_octs = 1 << 14  # 131,072 cells per call (a 64x64x32 Morton block)
        cells = 0
        tc0 = time.perf_counter()
        calls = 0
        while True:
            om.godunov_fine_sample(p, level, float(dts[-1]), dx, sample_octs)
            cells += sample_octs * 8
            calls += 1
            if time.perf_counter() - tc0 >= args.cpu_seconds:
                break
        tcpu = time.perf_counter() - tc0
        cpu = {
            "value": cells / tcpu,
            "unit": "cell-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{calls} godunov_fine calls of the oracle (single-threaded C++ restatement, g++ -O3, "
                      f"no FMA) on the first {sample_octs} octs ({sample_octs * 8} leaf cells) of the same "
                      f"level-{level} state, {tcpu:.1f}s on {cpu_model()}",
        }
    else:
        s.close()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"single-level roofline bench: {nleaf} leaf cells ({nleaf // 8} octs) at reference "
                        f"level {level}, periodic octree, lognormal rho/p seed {args.seed}, gamma=5/3, "
                        f"{args.riemann.upper()}, slope_type {args.slope}, FP64 plain form",
            "cells": int(nleaf),
            "level": level,
            "precision": "fp64",
            "riemann": args.riemann,
            "slope_type": args.slope,
            "kernel_ms": {"godunov": g_ms / max(g_n, 1), "courant": c_ms / max(args.steps, 1),
                          "set_unew": u_ms / max(args.steps, 1), "set_uold": o_ms / max(args.steps, 1)},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
