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
import math
import os
import platform
import subprocess
import sys
import time
from ctypes import byref, c_double as C_double, c_int as C_int, c_longlong as C_longlong

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ramses-cpp_amd"))

METRIC = "hydro cell-updates/s on 1 MI355X, FP64 vs mixed+batched; % of HBM roofline"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 80  # read uold (5 x FP64) + write unew (5 x FP64)

# Counter figures of the Godunov kernel from rocprofv3 runs of this script
# with --variant (profiles/), keyed by (level, riemann, slope_type, kernel
# variant, form), for a launch with the next step's Courant candidates:
#   traffic  HBM bytes per launch: FETCH_SIZE (one --pmc run) scaled by the
#            ratio it reads in the same run for k_set_unew_oct_lanes, a copy
#            of known size with the Godunov kernel's lane mapping (x1.998),
#            plus WRITE_SIZE (another run) as read: on a call's last step,
#            whose only writes are the 671,088,640 B of unew, it reads those
#            bytes to within 0.03 % (the copy's own write count runs 2-5 %
#            high and varies between runs);
#   busy     VALU utilisation (SQ_ACTIVE_INST_VALU x 4 / SIMDs / busy cycles);
#   insts_per_cell  VALU instructions per cell, SQ_INSTS_VALU / SQ_WAVES x
#            the waves per 64 cells (1, or 3 for variant 6).
# Counters cannot be read while timing.  Each entry records the hash of the
# kernel sources it was measured on (KERNEL_SOURCES); a build whose sources
# hash differently reports null instead of a stale figure.
KERNEL_SOURCES = ["ramses-cpp_amd/csrc/rh_hydro.hip", "ramses-cpp_amd/csrc/rh_physics.h",
                  "ramses-cpp_amd/csrc/rh_kernels.h", "ramses-cpp_amd/Makefile"]
PROFILED = {
    (9, "hllc", 1, 2, "fp64"): {"sources": "aaea3b054087f57a", "traffic": 2172901455,
                                "valu": {"busy": 0.761, "insts_per_cell": 5941.0, "courant_epilogue": True}},
    (9, "hllc", 1, 6, "fp64"): {"sources": "aaea3b054087f57a", "traffic": 2175126472,
                                "valu": {"busy": 0.812, "insts_per_cell": 6237.0, "courant_epilogue": True}},
    (9, "hllc", 1, 2, "mixed"): {"sources": "aaea3b054087f57a", "traffic": 2197427939,
                                 "valu": {"busy": 0.849, "insts_per_cell": 3958.2, "courant_epilogue": True}},
    (9, "hllc", 1, 6, "mixed"): {"sources": "aaea3b054087f57a", "traffic": 2178278029,
                                 "valu": {"busy": 0.685, "insts_per_cell": 4291.2, "courant_epilogue": True}},
}


def kernel_source_hash() -> str:
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profiled(level, riemann, slope, variant, form, what):
    """A counter figure of this workload and kernel build, or None."""
    e = PROFILED.get((level, riemann, slope, variant, form))
    if e is None or e["sources"] != kernel_source_hash():
        return None
    return e[what]


def reference_arithmetic(rate: float, riemann: str, slope: int):
    """The reference algorithm's FP64 operation rate when the Godunov kernel
    updates `rate` cells/s, against MI355X's FP64 peak (which counts an FMA as
    two operations; with fp-contract off every operation is its own
    instruction)."""
    ops = REFERENCE_FP64_OPS_PER_CELL.get((riemann, slope))
    if ops is None:
        return None
    tflops = rate * ops / 1e12
    return {"fp64_ops_per_cell_update": ops, "tflops": tflops, "peak_tflops": FP64_PEAK_TFLOPS,
            "frac": tflops / FP64_PEAK_TFLOPS, "counted_by": "oracle/flopcount.cpp"}


PROFILE_NOTE = "profiles/godunov_level9_256cubed_hllc_minmod.md (rocprofv3 counter runs of this script)"

# FP64 operations (add/sub, mul, div, sqrt) of the reference's godunov_fine per
# cell-update on this workload's state, counted exactly by running the
# oracle with a counting floating-point type (oracle/flopcount.cpp,
# tests/test_oracle_flops.py), keyed by (riemann, slope_type).
REFERENCE_FP64_OPS_PER_CELL = {("hllc", 1): 2412.6, ("hllc", 2): 2421.8, ("hll", 1): 2502.7, ("hll", 2): 2512.6,
                               ("llf", 1): 2412.7, ("llf", 2): 2422.6}
FP64_PEAK_TFLOPS = 78.6


def finite_or_null(x):
    """x with every non-finite float replaced by None (strict JSON)."""
    if isinstance(x, dict):
        return {k: finite_or_null(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [finite_or_null(v) for v in x]
    if isinstance(x, float) and not math.isfinite(x):
        return None
    return x


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def state_errors(A, B, gamma, np):
    """L1 (mean) and Linf relative errors of rho, p and |v| of conservative
    states A against B ([5, ncells] each); |v| relative to B's largest."""
    out = {}
    va, vb = A[1:4] / A[0], B[1:4] / B[0]
    pa = (gamma - 1) * (A[4] - 0.5 * A[0] * (va ** 2).sum(0))
    pb = (gamma - 1) * (B[4] - 0.5 * B[0] * (vb ** 2).sum(0))
    sa, sb = np.sqrt((va ** 2).sum(0)), np.sqrt((vb ** 2).sum(0))
    for name, x, y, scale in (("rho", A[0], B[0], np.abs(B[0])), ("p", pa, pb, np.abs(pb)),
                              ("v", sa, sb, np.full_like(sb, max(sb.max(), 1e-300)))):
        r = np.abs(x - y) / scale
        out[name] = {"L1": float(r.mean()), "Linf": float(r.max())}
    return out


def measure_ensemble(args, rhydro, np):
    """The batched form on the box ensemble: `boxes` independent periodic
    boxes of 16-64 cells a side, shock tubes and Sedov blasts, each with its
    own level, cell width and time step, advanced together (rh_run_ensemble),
    in the mixed and FP64 forms; then the same boxes one after another through
    rh_run_uniform (the unbatched alternative), whose final states must equal
    the batched ones."""
    gamma, cf = 1.4, 0.8
    boxes = rhydro.box_ensemble(args.boxes, seed=args.seed, gamma=gamma)
    nsedov = sum(b.kind == "sedov" for b in boxes)
    meshes = [b.mesh for b in boxes]
    levels = [b.level for b in boxes]
    dxs = [b.dx for b in boxes]
    u0 = [m.uold.copy() for m in meshes]
    sides = sorted(b.nx * 2 ** (b.level - 1) for b in boxes)
    warm = max(args.warmup, 1)
    out = {}
    finals = {}
    box_dts = {}
    for tag, prec in (("mixed", rhydro.RH_PREC_MIXED), ("fp64", rhydro.RH_PREC_FP64)):
        for m, u in zip(meshes, u0):
            m.uold[:] = u
            m.unew[:] = u
        s = rhydro.Solver(gamma=gamma, slope_type=1, riemann="hllc", precision=prec)
        s.bind_ensemble(meshes, rhydro.RH_DEVICE_RESIDENT)
        nleaf = s.ensemble_leaf_cells(levels)
        dts_warm = s.run_ensemble(levels, dxs, cf, warm)
        s.timers(True)
        s.timers_reset()
        t0 = time.perf_counter()
        dts_timed = s.run_ensemble(levels, dxs, cf, args.steps)  # synchronises before returning
        wall = time.perf_counter() - t0
        box_dts[tag] = np.concatenate([dts_warm, dts_timed])
        k = {name: s.timer(tk) for name, tk in (("godunov", rhydro.TK_GODUNOV), ("courant", rhydro.TK_COURANT),
                                                ("set_unew", rhydro.TK_SET_UNEW), ("set_uold", rhydro.TK_SET_UOLD),
                                                ("step_join", rhydro.TK_STEP_JOIN))}
        s.timers(False)
        s.sync_to_host()
        finals[tag] = [m.uold.copy() for m in meshes]
        # As many steps again without the timers' events, continuing from
        # the state: eager launches, then replayed from a graph, each after
        # an untimed call of the same length (which also captures the graph).
        untimed = {}
        for graphs in (False, True):
            s.graphs(graphs)
            s.run_ensemble(levels, dxs, cf, args.steps)
            t1 = time.perf_counter()
            s.run_ensemble(levels, dxs, cf, args.steps)
            untimed["graph" if graphs else "eager"] = (time.perf_counter() - t1) / args.steps * 1e3
        s.close()
        god_s = k["godunov"][0] / max(k["godunov"][1], 1) / 1e3
        out[tag] = {
            "value": nleaf * args.steps / wall,
            "ms_per_step": wall / args.steps * 1e3,
            "kernel_ms": {n: v[0] / max(v[1], 1) for n, v in k.items()},
            "roofline_frac": BYTES_PER_CELL_UPDATE * nleaf / god_s / 1e9 / HBM_PEAK_GBS,
            "untimed_ms_per_step": untimed,
            "value_graph": nleaf / (untimed["graph"] * 1e-3),
        }
        print(f"[bench] ensemble {tag}: {nleaf} cells, {args.steps} steps in {wall:.4f}s, godunov "
              f"{god_s * 1e3:.3f} ms", file=sys.stderr, flush=True)
    # Unbatched: every box alone, mixed form, same warmup and steps.
    s = rhydro.Solver(gamma=gamma, slope_type=1, riemann="hllc", precision=rhydro.RH_PREC_MIXED)
    t_loop = 0.0
    same = True
    for b, (m, u) in enumerate(zip(meshes, u0)):
        m.uold[:] = u
        m.unew[:] = u
        s.bind(m, rhydro.RH_DEVICE_RESIDENT)
        s.run_uniform(levels[b], dxs[b], cf, warm)
        t0 = time.perf_counter()
        s.run_uniform(levels[b], dxs[b], cf, args.steps)
        t_loop += time.perf_counter() - t0
        s.sync_to_host()
        same = same and bool(np.array_equal(m.uold, finals["mixed"][b], equal_nan=True))
    s.close()
    unbatched = nleaf * args.steps / t_loop
    print(f"[bench] unbatched mixed loop: {t_loop:.4f}s, identical={same}", file=sys.stderr, flush=True)
    # Per-box error of the mixed form against the FP64 form (bit-identical to
    # the oracle) after warmup + steps, each form with its own time steps.
    # A Sedov box has blown up when its FP64 run (the reference scheme's)
    # took a time step below 1e-6 of its first one: on this initial state the
    # scheme can produce negative densities and pressures and speeds of 1e10
    # and more within its first steps (the reference's own run of the 64^3
    # blast logs a second time step of 5.6e-15 after 1.8e-4), after which the
    # floors leave a state with time steps near 1e-14.  The two forms then
    # follow a blown-up solution, and their difference says nothing about
    # precision, so those boxes are summarised apart.
    per_box, group = [], []
    for b, m in enumerate(meshes):
        cells = m.cells_of_level(levels[b], leaf_only=True)
        A = finals["mixed"][b].reshape(5, -1)[:, cells]
        B = finals["fp64"][b].reshape(5, -1)[:, cells]
        per_box.append(state_errors(A, B, gamma, np))
        d64 = box_dts["fp64"][:, b]
        collapsed = boxes[b].kind == "sedov" and float(d64.min()) < 1e-6 * float(d64[0])
        group.append("sedov_collapsed" if collapsed else boxes[b].kind)

    def summary(members):
        res = {"boxes": len(members)}
        if not members:
            return res
        for name in ("rho", "p", "v"):
            for norm in ("L1", "Linf"):
                vals = np.array([per_box[b][name][norm] for b in members])
                ok = np.isfinite(vals)
                if not ok.any():
                    res[f"{name}_{norm}"] = None
                    continue
                worst = int(np.flatnonzero(ok)[int(vals[ok].argmax())])
                res[f"{name}_{norm}"] = {"max": float(vals[ok].max()), "median": float(np.median(vals[ok])),
                                         "worst_box": int(members[worst]), "nonfinite_boxes": int((~ok).sum())}
        return res

    err_summary = {g: summary([b for b in range(len(boxes)) if group[b] == g]) for g in ("sod", "sedov")}
    blown = [b for b in range(len(boxes)) if group[b] == "sedov_collapsed"]
    err_summary["sedov_blown_up"] = {"boxes": len(blown), "box_ids": blown}
    return {
        "workload": f"box ensemble: {args.boxes} periodic cubic boxes, sides {sides[0]}-{sides[-1]} cells "
                    f"(log-uniform, nx*2^(L-1) with nx in 1,3,5,7; seed {args.seed}), {nleaf} leaf cells; "
                    f"{args.boxes - nsedov} shock tubes (random plane normal, rho and p in [0.1, 1] per side) and "
                    f"{nsedov} Sedov blasts (energy in [0.1, 1] in the 8 central cells, rho=1, p=1e-5); gamma=1.4, "
                    "HLLC, slope_type 1, own Courant time step per box",
        "boxes_sod": args.boxes - nsedov,
        "boxes_sedov": nsedov,
        "cells": int(nleaf),
        "value_mixed": out["mixed"]["value"],
        "value_fp64": out["fp64"]["value"],
        "mixed": out["mixed"],
        "fp64": out["fp64"],
        "unbatched_mixed_value": unbatched,
        "batched_speedup": out["mixed"]["value"] / unbatched,
        "batched_equals_unbatched": same,
        "mixed_error_vs_fp64_per_box": err_summary,
        "error_note": "relative errors of the mixed form against the FP64 form on each box's leaf cells after "
                      "warmup + steps, each form with its own time steps; max and median over the shock tubes, "
                      "the Sedov boxes; the Sedov boxes whose FP64 run (the reference scheme's) has blown up "
                      "(a time step below 1e-6 of its first) are only listed",
    }


def measure_amr(args, rhydro, np, which, precision=None):
    """Adaptive runs with the state on the GPU (rh_amr_*): the AMR Sedov blast
    (reference levels 7-10, err_grad_d = err_grad_p = 0.1) or the
    Kelvin-Helmholtz slab (levels 8-11), timed over whole coarse steps
    (hydro, refinement criteria, tree edits, prolongation, restriction), in
    the FP64 form unless `precision` says otherwise.  Cell-updates = leaf
    cells of every godunov_fine call.  For the Sedov run the first two coarse
    steps are also run through the oracle on one host core and compared value
    for value."""
    precision = rhydro.RH_PREC_FP64 if precision is None else precision
    if which == "sedov":
        kw = dict(levelmin=7, levelmax=10, ngridmax=300000, err_grad_d=0.1, err_grad_p=0.1,
                  regions=rhydro.sedov_regions(7))
        steps = args.amr_steps
        name = ("AMR Sedov blast: 64^3 base (reference level 7) refined to level 10 on density and pressure "
                "gradients (err_grad 0.1), subcycled, gamma=1.4, HLLC, slope_type 1, interpol_type 1")
    else:
        kw = dict(levelmin=8, levelmax=11, ngridmax=4500000, err_grad_d=0.2, ic="kh", kh_amp=0.01)
        steps = args.kh_steps
        name = ("AMR Kelvin-Helmholtz slab: 128^3 base (reference level 8) refined to level 11 on density "
                "gradients (err_grad_d 0.2), subcycled, gamma=1.4, HLLC, slope_type 1")
    common = dict(gamma=1.4, slope_type=1, riemann="hllc", interpol_type=1, interpol_var=0)
    L = rhydro.lib()
    tc0 = time.perf_counter()
    run = rhydro.AmrRun(**kw, **common, precision=precision)
    t_init = time.perf_counter() - tc0
    # Kernel timers put two events around every launch, which the Sedov run's
    # hundreds of short launches per coarse step would feel: there the timed
    # run has them off and a second, identical run reads the Godunov time.
    timed_run = which != "sedov"
    ctx = run.solver_context()
    tune = None
    if which != "sedov":
        # Godunov kernel variant picked on the finest level (the bulk of the
        # cell-updates), after the initial refinement.
        lmax = kw["levelmax"]
        dxl = 1.0 / 2 ** lmax
        dt = C_double()
        if L.rh_courant(ctx, lmax, dxl, 1.4, 0.8, byref(dt)) == 0:
            bv = C_int()
            msv = (C_double * rhydro.RH_KERNEL_VARIANTS)()
            if L.rh_autotune(ctx, lmax, dt.value, dxl, 2, byref(bv), msv) == 0:
                tune = {"level": lmax, "chosen_variant": bv.value, "godunov_ms_by_variant": list(msv)}
    L.rh_timers_enable(ctx, 1 if timed_run else 0)
    L.rh_timers_reset(ctx)
    _, _, cu0 = run.stats()
    t0 = time.perf_counter()
    ph0 = run.phase_ms()
    run.step(steps)
    wall = time.perf_counter() - t0
    ph1 = run.phase_ms()
    t_sim, nstep, cu1 = run.stats()
    ms, n = C_double(), C_longlong()
    st = run.state()
    if timed_run:
        L.rh_timers_read(ctx, rhydro.TK_GODUNOV, byref(ms), byref(n))
    else:
        again = rhydro.AmrRun(**kw, **common, precision=precision)
        L.rh_timers_enable(again.solver_context(), 1)
        L.rh_timers_reset(again.solver_context())
        again.step(steps)
        L.rh_timers_read(again.solver_context(), rhydro.TK_GODUNOV, byref(ms), byref(n))
        again.close()
    run.close()
    out = {"workload": name, "precision": "mixed" if precision == rhydro.RH_PREC_MIXED else "fp64",
           "coarse_steps": steps, "fine_steps": nstep, "cell_updates": int(cu1 - cu0),
           "value": (cu1 - cu0) / wall, "wall_s": wall, "init_s": t_init, "godunov_ms_total": ms.value,
           "godunov_launches": n.value, "octs_per_level": st["numbl"].tolist(), "autotune": tune,
           "host_phase_ms": {k: ph1[k] - ph0[k] for k in ph1}, "_state": st}
    print(f"[bench] amr {which}: {steps} coarse steps, {cu1 - cu0} cell-updates in {wall:.3f}s "
          f"(godunov {ms.value:.1f} ms), lists {st['numbl'].tolist()}", file=sys.stderr, flush=True)
    if not args.no_cpu and which == "sedov" and precision == rhydro.RH_PREC_FP64:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        okw = dict(kw, riemann=rhydro.resolve_riemann("hllc"), gamma=1.4, slope_type=1, interpol_type=1)
        o = oracle.OracleAmr(**okw)
        g = rhydro.AmrRun(**kw, **common)
        ncmp = 2
        tcpu0 = time.perf_counter()
        o.step(ncmp)
        tcpu = time.perf_counter() - tcpu0
        g.step(ncmp)
        so, sg = o.state(), g.state()
        same = all(np.array_equal(so[k], sg[k], equal_nan=True)
                   for k in ("numbl", "son", "father", "nbor", "next", "uold", "unew"))
        out.update({"cpu_value": so["cell_updates"] / tcpu, "cpu_kind": "port", "cpu_cores": 1,
                    "cpu_sample": f"the first {ncmp} coarse steps through the oracle",
                    "identical_to_oracle_after_2_steps": bool(same)})
        g.close()
        o.close()
        print(f"[bench] amr {which}: oracle {ncmp} steps {tcpu:.1f}s identical={same}", file=sys.stderr, flush=True)
    return out


def amr_errors(ref, other, gamma, np):
    """Mixed against FP64 after an adaptive run: the refinement decisions of
    the two forms can differ, so the errors are taken on the leaf cells only
    when both trees are identical."""
    same_tree = all(np.array_equal(ref[k], other[k]) for k in ("numbl", "headl", "next", "son", "father"))
    out = {"mesh_identical_to_fp64": bool(same_tree)}
    if not same_tree:
        out["cells_refined_differently"] = int((ref["son"] != other["son"]).sum())
        return out
    leaf = np.flatnonzero(ref["son"][ref["ncoarse"]:] == 0) + ref["ncoarse"]
    inlist = np.zeros(ref["ncell"], bool)
    for L in range(1, ref["nlevelmax"] + 1):
        g, seq = int(ref["headl"][L]), []
        while g > 0:
            seq.append(g)
            g = int(ref["next"][g - 1])
        if seq:
            gi = np.array(seq)
            for ic in range(8):
                inlist[ref["ncoarse"] + ic * ref["ngridmax"] + gi - 1] = True
    cells = leaf[inlist[leaf]]
    A = other["uold"].reshape(5, -1)[:, cells]
    B = ref["uold"].reshape(5, -1)[:, cells]
    out["error_vs_fp64"] = state_errors(A, B, gamma, np)
    out["error_note"] = "relative errors of the mixed form against the FP64 form on all leaf cells after the run"
    return out


def measure_sedov64(args, rhydro, np):
    """The 64^3 Sedov blast (reference level 7, 262,144 cells; p = 1e-5
    outside, the unit energy in the 8 central cells), gamma = 1.4, HLLC,
    MinMod, 20 steps of the FP64 form from the initial state, timed on the
    GPU and, unless --no-cpu, through the oracle on one host core; the two
    final states are compared value for value."""
    level, nsteps, cf = 7, 20, 0.8
    mesh = rhydro.uniform_mesh(level)
    rhydro.init_regions(mesh, rhydro.sedov_regions(level), 1.4)
    u0 = mesh.uold.copy()
    dx = 1.0 / 2 ** level
    s = rhydro.Solver(gamma=1.4, slope_type=1, riemann="hllc")
    s.bind(mesh, rhydro.RH_DEVICE_RESIDENT)
    nleaf = s.leaf_cells(level)
    s.run_uniform(level, dx, cf, 1)  # first use: tables and kernels loaded
    mesh.uold[:] = u0
    mesh.unew[:] = u0
    s.bind(mesh, rhydro.RH_DEVICE_RESIDENT)
    # 20 steps from the initial state: once eagerly, then (graphs on, the
    # default) a first call that captures the launch graph and a timed one
    # that replays it.
    walls = {}
    for graphs in (False, True):
        s.graphs(graphs)
        for rep in range(2 if graphs else 1):
            mesh.uold[:] = u0
            mesh.unew[:] = u0
            s.sync_to_device()
            t0 = time.perf_counter()
            dts = s.run_uniform(level, dx, cf, nsteps)
            walls["graph" if graphs else "eager"] = time.perf_counter() - t0
    wall = walls["graph"]
    s.sync_to_host()
    s.close()
    out = {"workload": "Sedov blast 64^3 (reference level 7), gamma=1.4, HLLC, slope_type 1, 20 steps, FP64 form",
           "cells": int(nleaf), "steps": nsteps, "gpu_value": nleaf * nsteps / wall, "gpu_wall_s": wall,
           "gpu_value_eager": nleaf * nsteps / walls["eager"], "dt_first": float(dts[0]), "dt_last": float(dts[-1])}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        ref = rhydro.uniform_mesh(level)
        ref.uold[:] = u0
        ref.unew[:] = u0
        om = oracle.OracleMesh(ref)
        tc = time.perf_counter()
        dto = om.run_uniform(oracle.params(1.4, 1, 2), level, dx, cf, nsteps)
        tcpu = time.perf_counter() - tc
        out.update({"cpu_value": nleaf * nsteps / tcpu, "cpu_wall_s": tcpu, "cpu_kind": "port", "cpu_cores": 1,
                    "identical_to_oracle": bool(np.array_equal(dts, dto, equal_nan=True)
                                                and np.array_equal(mesh.uold, ref.uold, equal_nan=True))})
    print(f"[bench] sedov64: gpu {wall:.4f}s" + (f", cpu {out['cpu_wall_s']:.1f}s, identical "
                                                 f"{out['identical_to_oracle']}" if "cpu_wall_s" in out else ""),
          file=sys.stderr, flush=True)
    return out


def measure_dropin(args, rhydro, np, which):
    """The drop-in path as the HydroSolver adapter drives it
    (ramses-cpp_amd/adapter/PatchHydroSolver.cpp), one amr_step of a uniform
    level = compute_courant_step, set_unew, godunov_fine, set_uold
    (Simulation.cpp:550,564,589,614), each through its own C-ABI call:
      host-coherent (what an adaptive run needs): every call copies uold and
        unew to the GPU and the array it wrote back, and the adapter first
        compares the topology arrays with its copy from the last binding;
      device-resident (the adapter's mode for uniform runs): the state stays
        on the GPU and uold comes back once per step, after set_uold.
    Both leave the same state (compared at the end).  `which`: "c3" is the
    256^3 lognormal workload of the main measurement, "c1" the 64^3 Sedov
    blast."""
    if which == "c3":
        level, gamma, slope, nsteps = 9, 5.0 / 3.0, 1, 3
        mesh = rhydro.uniform_mesh(level)
        rhydro.init_lognormal(mesh, level, args.seed, gamma, 0.5, 0.5)
        name = "256^3 lognormal workload (reference level 9), gamma=5/3, HLLC, slope_type 1"
    else:
        level, gamma, slope, nsteps = 7, 1.4, 1, 10
        mesh = rhydro.uniform_mesh(level)
        rhydro.init_regions(mesh, rhydro.sedov_regions(level), gamma)
        name = "Sedov blast 64^3 (reference level 7), gamma=1.4, HLLC, slope_type 1"
    dx = 1.0 / 2 ** level
    u0 = mesh.uold.copy()
    out = {"workload": name + ", one C-ABI call per HydroSolver virtual, FP64 form", "steps": nsteps}
    finals = {}
    for mode_name, mode in (("host_coherent", rhydro.RH_HOST_COHERENT), ("device_resident", rhydro.RH_DEVICE_RESIDENT)):
        mesh.uold[:] = u0
        mesh.unew[:] = u0
        s = rhydro.Solver(gamma=gamma, slope_type=slope, riemann="hllc")
        s.bind(mesh, mode)
        nleaf = s.leaf_cells(level)
        topo = [mesh.arrays[k] for k in ("father", "son", "nbor", "next", "headl")]
        saved = [a.copy() for a in topo]
        calls = 0.0
        t0 = time.perf_counter()
        for k in range(nsteps):
            for _ in range(4 if mode == rhydro.RH_HOST_COHERENT else 0):
                # the adapter's per-call topology test (memcmp of the saved arrays)
                assert all(np.array_equal(a, b) for a, b in zip(topo, saved))
            tc = time.perf_counter()
            dt = s.courant(level, dx, gamma, 0.8)
            s.set_unew(level)
            s.godunov_fine(level, dt, dx)
            s.set_uold(level)
            if mode == rhydro.RH_DEVICE_RESIDENT:
                s.sync_to_host(rhydro.RH_UOLD)
            calls += time.perf_counter() - tc
        wall = time.perf_counter() - t0
        s.close()
        finals[mode_name] = mesh.uold.copy()
        out[mode_name] = {"value": nleaf * nsteps / wall, "ms_per_step": wall / nsteps * 1e3,
                          "topology_check_ms_per_step": (wall - calls) / nsteps * 1e3}
        print(f"[bench] drop-in {which} {mode_name}: {wall / nsteps * 1e3:.2f} ms/step", file=sys.stderr, flush=True)
    out["cells"] = int(nleaf)
    out["modes_identical"] = bool(np.array_equal(finals["host_coherent"], finals["device_resident"], equal_nan=True))
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--level", type=int, default=9, help="reference level list (9 -> 256^3 cells)")
    ap.add_argument("--riemann", default="hllc")
    ap.add_argument("--slope", type=int, default=1)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", type=int, default=None,
                    help="Godunov kernel variant of the main workload (default: autotuned on the level)")
    ap.add_argument("--fp64-only", action="store_true", help="skip the mixed-precision form")
    ap.add_argument("--boxes", type=int, default=96, help="boxes of the shock-tube and Sedov ensemble (0: skip it)")
    ap.add_argument("--no-variants", action="store_true", help="skip the HLL / slope-2 variants of the workload")
    ap.add_argument("--no-sedov", action="store_true", help="skip the 64^3 Sedov run")
    ap.add_argument("--no-dropin", action="store_true", help="skip the drop-in (per-virtual C-ABI call) measurement")
    ap.add_argument("--only", choices=["ensemble", "amr_sedov", "amr_kh", "sedov64"],
                    help="run one secondary workload alone and print its record (for profiling)")
    ap.add_argument("--amr-steps", type=int, default=50, help="coarse steps of the AMR Sedov run (0: skip)")
    ap.add_argument("--kh-steps", type=int, default=5, help="coarse steps of the AMR Kelvin-Helmholtz run (0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (one GPU per run here)",
              file=sys.stderr)
        return 2

    import numpy as np

    import rhydro

    if args.only:
        if args.only == "ensemble":
            rec = measure_ensemble(args, rhydro, np)
        elif args.only == "sedov64":
            rec = measure_sedov64(args, rhydro, np)
        else:
            rec = measure_amr(args, rhydro, np, "sedov" if args.only == "amr_sedov" else "kh")
            rec.pop("_state", None)
        print(json.dumps(finite_or_null({"only": args.only, "record": rec}), allow_nan=False))
        return 0

    gamma = 5.0 / 3.0
    level = args.level
    t_setup = time.time()
    mesh = rhydro.uniform_mesh(level)
    rhydro.init_lognormal(mesh, level, args.seed, gamma, 0.5, 0.5)
    print(f"[bench] mesh level {level}: {mesh.ngridmax} grids, {mesh.ncell} cells, setup "
          f"{time.time() - t_setup:.1f}s", file=sys.stderr, flush=True)

    dx = 1.0 / (2 ** level)  # Simulation.cpp:548, boxlen/(nx*2^ilevel)
    cf = 0.8
    u0 = mesh.uold.copy()

    def measure(precision, riemann=args.riemann, slope=args.slope):
        """Warmup, then exactly `steps` timed steps of one precision form."""
        mesh.uold[:] = u0
        mesh.unew[:] = u0
        s = rhydro.Solver(gamma=gamma, slope_type=slope, riemann=riemann, precision=precision,
                          variant=args.variant)
        s.bind(mesh, rhydro.RH_DEVICE_RESIDENT)
        nleaf = s.leaf_cells(level)
        # Pick the faster Godunov kernel variant on this level (results are
        # identical for every variant; unew is restored by the tuner) unless
        # one is given, then the warmup steps right before the timed ones, so
        # that no host work leaves the GPU idle in between.
        if args.variant is None:
            best, tune_ms = s.autotune(level, s.courant(level, dx, gamma, cf), dx, reps=2)
        else:
            best, tune_ms = args.variant, np.full(rhydro.RH_KERNEL_VARIANTS, np.nan)
        s.run_uniform(level, dx, cf, max(args.warmup, 1))  # ends with a stream synchronisation
        s.timers(True)
        s.timers_reset()
        t0 = time.perf_counter()
        dts = s.run_uniform(level, dx, cf, args.steps)  # synchronises before returning
        wall = time.perf_counter() - t0
        k = {name: s.timer(tk) for name, tk in (("godunov", rhydro.TK_GODUNOV), ("courant", rhydro.TK_COURANT),
                                                ("set_unew", rhydro.TK_SET_UNEW), ("set_uold", rhydro.TK_SET_UOLD),
                                                ("step_join", rhydro.TK_STEP_JOIN))}
        s.timers(False)
        s.sync_to_host()
        final = mesh.uold.copy()
        # A copy of known size (the level's 671 MB of state, uold -> unew,
        # equal already) with the Godunov kernel's lane mapping: its byte
        # counters in a profiler run calibrate the kernel's.
        s.set_unew_oct_lanes(level)
        # As many steps again without the timers' events, continuing from
        # the state: eager launches, then replayed from a graph, each after
        # an untimed call of the same length (which warms the GPU up again
        # and captures the graph).
        untimed = {}
        for graphs in (False, True):
            s.graphs(graphs)
            s.run_uniform(level, dx, cf, args.steps)
            t1 = time.perf_counter()
            s.run_uniform(level, dx, cf, args.steps)
            untimed["graph" if graphs else "eager"] = (time.perf_counter() - t1) / args.steps * 1e3
        s.close()
        god_avg_s = k["godunov"][0] / max(k["godunov"][1], 1) / 1e3
        achieved = BYTES_PER_CELL_UPDATE * nleaf / god_avg_s / 1e9
        tag = "fp64" if precision == rhydro.RH_PREC_FP64 else "mixed"
        print(f"[bench] {tag}: {args.steps} steps in {wall:.4f}s, godunov {god_avg_s * 1e3:.3f} ms, untimed "
              f"eager {untimed['eager']:.3f} / graph {untimed['graph']:.3f} ms per step", file=sys.stderr, flush=True)
        return dict(nleaf=nleaf, wall=wall, dts=dts, achieved=achieved, final=final,
                    kernel_ms={n: v[0] / max(v[1], 1) for n, v in k.items()},
                    autotune={"chosen_variant": best, "given": args.variant is not None,
                              "godunov_ms_by_variant": tune_ms.tolist()},
                    untimed_ms_per_step=untimed)

    fp64 = measure(rhydro.RH_PREC_FP64)
    mixed = measure(rhydro.RH_PREC_MIXED) if not args.fp64_only else None
    nleaf = fp64["nleaf"]

    def one_step(precision):
        mesh.uold[:] = u0
        mesh.unew[:] = u0
        s = rhydro.Solver(gamma=gamma, slope_type=args.slope, riemann=args.riemann, precision=precision)
        s.bind(mesh, rhydro.RH_DEVICE_RESIDENT)
        s.run_uniform(level, dx, cf, 1)
        s.sync_to_host()
        s.close()
        return mesh.uold.copy()

    def errors(a, b):
        """L1 (mean) and Linf relative errors of rho, p, |v| of state a against b."""
        cells = mesh.cells_of_level(level)
        return state_errors(a.reshape(5, -1)[:, cells], b.reshape(5, -1)[:, cells], gamma, np)

    mixed_info = None
    if mixed is not None:
        err1 = errors(one_step(rhydro.RH_PREC_MIXED), one_step(rhydro.RH_PREC_FP64))
        errn = errors(mixed["final"], fp64["final"])
        mixed_info = {
            "value": nleaf * args.steps / mixed["wall"],
            "ms_per_step": mixed["wall"] / args.steps * 1e3,
            "kernel_ms": mixed["kernel_ms"],
            "untimed_ms_per_step": mixed["untimed_ms_per_step"],
            "autotune": mixed["autotune"],
            "roofline_frac": mixed["achieved"] / HBM_PEAK_GBS,
            "traffic": profiled(level, args.riemann, args.slope, mixed["autotune"]["chosen_variant"], "mixed",
                                "traffic"),
            "valu": profiled(level, args.riemann, args.slope, mixed["autotune"]["chosen_variant"], "mixed", "valu"),
            "speedup_vs_fp64": fp64["wall"] / mixed["wall"],
            "error_vs_fp64_one_step": err1,
            "error_vs_fp64_after_steps": errn,
            "error_note": "relative errors of the mixed form against the FP64 form (bit-identical to the "
                          "reference restatement) on leaf cells; one step from identical input, and after "
                          f"{args.warmup}+{args.steps} steps with each form's own time steps",
        }

    # The other Riemann solver / slope combinations of the same workload:
    # godunov kernel time and roofline fraction per form.
    variants = []
    if not args.no_variants:
        for rv, sv in (("hllc", 1), ("hllc", 2), ("hll", 1), ("hll", 2)):
            if (rv, sv) == (args.riemann, args.slope):
                a, b = fp64, mixed
            else:
                a = measure(rhydro.RH_PREC_FP64, rv, sv)
                b = measure(rhydro.RH_PREC_MIXED, rv, sv) if not args.fp64_only else None
            row = {"riemann": rv, "slope_type": sv, "fp64_value": nleaf * args.steps / a["wall"],
                   "fp64_godunov_ms": a["kernel_ms"]["godunov"], "fp64_roofline_frac": a["achieved"] / HBM_PEAK_GBS}
            if b is not None:
                row.update({"mixed_value": nleaf * args.steps / b["wall"], "mixed_godunov_ms": b["kernel_ms"]["godunov"],
                            "mixed_roofline_frac": b["achieved"] / HBM_PEAK_GBS})
            variants.append(row)

    ens_info = measure_ensemble(args, rhydro, np) if args.boxes > 0 else None
    dropin = None if args.no_dropin else {w: measure_dropin(args, rhydro, np, w) for w in ("c3", "c1")}
    sedov_info = measure_sedov64(args, rhydro, np) if not args.no_sedov else None
    amr_sedov = measure_amr(args, rhydro, np, "sedov") if args.amr_steps > 0 else None
    amr_kh = measure_amr(args, rhydro, np, "kh") if args.kh_steps > 0 else None
    if amr_kh is not None:
        m = measure_amr(args, rhydro, np, "kh", rhydro.RH_PREC_MIXED)
        amr_kh["mixed"] = {k: m[k] for k in ("value", "wall_s", "cell_updates", "fine_steps", "godunov_ms_total",
                                             "host_phase_ms")}
        amr_kh["mixed"].update(amr_errors(amr_kh["_state"], m["_state"], 1.4, np))
    for rec in (amr_sedov, amr_kh):
        if rec is not None:
            rec.pop("_state", None)

    value = nleaf * args.steps / fp64["wall"]
    achieved = fp64["achieved"]
    g_ms = fp64["kernel_ms"]["godunov"]

    cpu = None
    if not args.no_cpu and args.cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        mesh.uold[:] = fp64["final"]
        mesh.unew[:] = fp64["final"]
        om = oracle.OracleMesh(mesh)
        p = oracle.params(gamma, args.slope, rhydro.resolve_riemann(args.riemann), 1, 0, 1.0)
        sample_octs = 1 << 14  # 131,072 cells per call (a 64x64x32 Morton block)
        cells = 0
        tc0 = time.perf_counter()
        calls = 0
        while True:
            om.godunov_fine_sample(p, level, float(fp64["dts"][-1]), dx, sample_octs)
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

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp64["wall"] / args.steps * 1e3,
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
            "kernel_ms": fp64["kernel_ms"],
            # The timed steps above launch eagerly with HIP events around each
            # kernel; the same steps without events, eager and from a graph.
            "untimed_ms_per_step": fp64["untimed_ms_per_step"],
            "autotune": fp64["autotune"],
            "valu": profiled(level, args.riemann, args.slope, fp64["autotune"]["chosen_variant"], "fp64", "valu"),
            "counters_from": PROFILE_NOTE,
            "kernel_sources_sha16": kernel_source_hash(),
            "reference_arithmetic": reference_arithmetic(nleaf / (g_ms * 1e-3), args.riemann, args.slope),
            "mixed": mixed_info,
            "variants": variants,
            "ensemble": ens_info,
            "drop_in": dropin,
            "sedov64": sedov_info,
            "amr_sedov": amr_sedov,
            "amr_kelvin_helmholtz": amr_kh,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled(level, args.riemann, args.slope, fp64["autotune"]["chosen_variant"], "fp64", "traffic"),
            # What the counters show limits the kernel (the roofline above is
            # the metric's HBM one).
            "limiter": "VALU issue (SQ counters; see config.valu)",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(finite_or_null(out), allow_nan=False))
    return 0


if __name__ == "__main__":
    sys.exit(main())
