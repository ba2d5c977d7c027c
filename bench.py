#!/usr/bin/env python3
"""Throughput of the fine-level Godunov hydro step on one MI355X.

Workload: uniform 256^3 supersonic decaying turbulence (gamma=5/3, rho=1, p=1,
solenoidal velocity with P(k)~k^-2 on 2<=k<=8 at rms Mach 5, seeded), MonCen
slopes, HLLC, Courant 0.5, periodic.  A step is one full level step: the
Courant dt reduction (courant_fine) and the Godunov update (godunov_fine with
set_unew/set_uold folded in), all on the device.  The state is resident in HBM
before the timed region starts.

Prints one JSON line: value = FP64 cell-updates/s of the whole step, the K
timed steps replayed as the captured hipGraph of a step pair; the mixed
precision (FP32 reconstruction/fluxes, FP64 accumulation) rate sits beside it,
with a 100-step precision study (L1 per conserved variable against FP64, mass
and energy drift); roofline = the Godunov kernel's algorithmic HBM bytes (80 B
per cell-update) over its mean launch time, measured with HIP events on the
library stream in a separate eager pass, with the kernel's VALU-issue ceiling
from the committed counters beside it; cpu_baseline = the CPU oracle on the
host cores on a bounded sample (all threads, and one).  Further legs: uniform
Sedov 128^3 GPU vs CPU, AMR Sedov levels 6-10, ensemble serving.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SIMDS = 1024          # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9      # peak engine clock
BYTES_PER_CELL = 80.0  # read uold 5x8 B + write unew 5x8 B


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # the first 20-30 Godunov launches after host-side set-up run up to 15% slower
    # (the GPU's clocks ramp up): warm-up steps of every leg run untimed first
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--ncell", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-sample", type=int, default=128, help="cells per dimension of the CPU sample")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--eager", action="store_true",
                    help="time the steps launch by launch instead of replaying the captured step-pair hipGraph")
    ap.add_argument("--sedov-steps", type=int, default=200, help="steps of the uniform 128^3 Sedov leg")
    ap.add_argument("--no-sedov", action="store_true")
    ap.add_argument("--study-steps", type=int, default=100, help="steps of the FP64-vs-mixed precision study")
    ap.add_argument("--serve-jobs", type=int, default=2000, help="jobs of the ensemble serving stream")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="offered load as a fraction of the measured FP64 batched capacity")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="arrival rate in jobs/s (default: serve-load times the measured FP64 capacity)")
    ap.add_argument("--serve-steps", type=int, default=100)
    ap.add_argument("--serve-depth", type=int, default=2, help="ticks the request loop keeps enqueued")
    ap.add_argument("--serve-trial-jobs", type=int, default=300,
                    help="jobs of the stream's start served per candidate brick shape to pick each precision's shape")
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--amr-steps", type=int, default=20)
    ap.add_argument("--amr-levels", type=str, default="6,10", help="levelmin,levelmax of the AMR Sedov case")
    ap.add_argument("--no-amr", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the library's default 8x6x8 brick shape (no autotuning)")
    ap.add_argument("--fp64-brick", type=str, default=None, help="brick shape bx,by,bz of the FP64 run (no tuning)")
    ap.add_argument("--mixed-brick", type=str, default=None, help="brick shape of the mixed run (no tuning)")
    ap.add_argument("--amr-brick", type=str, default=None, help="brick shape of every AMR level (no tuning)")
    ap.add_argument("--serve-brick", type=str, default=None, help="brick shape of the serving jobs (no tuning)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one plain device copy of this many MiB first (calibrates HBM byte counters)")
    return ap.parse_args()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_rate(mesh, cells, p, threads, seconds):
    """Full steps (courant_fine + set_unew + godunov_fine + set_uold) of the CPU
    oracle (C restatement, OpenMP over octs) on `threads` host threads for
    about `seconds`: cell-updates/s."""
    from oracle import oracle as O
    from tests.helpers import oracle_params, oracle_uniform

    level = mesh.level
    om = oracle_uniform(mesh, mesh.from_cells(cells))
    po = oracle_params(p)
    octmap = (mesh.octmap + 1).astype(np.int32)
    O.lib()

    def one_step():
        c = om.courant_fine(level, po, threads)
        om.set_unew(level)
        om.godunov_uniform_omp(level, c[4], po, octmap, threads)
        om.set_uold(level)

    t0 = time.perf_counter()
    one_step()
    t1 = time.perf_counter()
    nsteps = max(1, min(200, int(seconds / max(t1 - t0, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(nsteps):
        one_step()
    el = time.perf_counter() - t0
    return {"cell_updates_per_s": nsteps * mesh.noct * 8 / el, "threads": threads, "steps": nsteps, "seconds": el}


def cpu_baseline(args, p):
    """The CPU oracle on a bounded sample of the headline problem (the same
    turbulence at cpu_sample^3 cells), on every host thread and on one."""
    from ramses2_amd import UniformMesh, ic

    n = args.cpu_sample
    mesh = UniformMesh(3, int(np.log2(n)))
    cells = ic.turbulence_3d(n, seed=args.seed)
    threads = len(os.sched_getaffinity(0))
    allt = cpu_rate(mesh, cells, p, threads, args.cpu_seconds)
    one = cpu_rate(mesh, cells, p, 1, args.cpu_seconds / 3)
    return {"value": allt["cell_updates_per_s"], "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "one_thread": one["cell_updates_per_s"],
            "sample": "CPU oracle (oracle/ramses_oracle.c, the reference's per-oct algorithm, OpenMP over octs) "
                      "on the same turbulence problem at %d^3 cells: %d full steps (courant_fine + godunov_fine) "
                      "in %.1f s on %d threads; %d steps in %.1f s on 1 thread (one_thread); host CPU: %s"
                      % (n, allt["steps"], allt["seconds"], threads, one["steps"], one["seconds"], cpu_model())}


def valu_ceiling(shape, ncells, kernel_ms):
    """The FP64 kernel's VALU-issue ceiling from the committed counter summary
    (when it was measured on this brick shape): every SIMD issuing one wave64
    VALU instruction per 4 cycles at the 2.4 GHz peak engine clock (1024
    SIMDs), divided by the counted VALU wave-instructions per cell-update."""
    path = os.path.join(ROOT, "profiles", "godunov_fp64_turb256_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("brick_shape") != "%dx%dx%d" % shape:
        return None
    ipc = d["valu_wave_instructions_per_cell"]
    ceiling = SIMDS * CLOCK_HZ / 4.0 / ipc
    achieved = ncells / (kernel_ms * 1e-3)
    return {"valu_wave_instructions_per_cell": ipc, "ceiling_cell_updates_per_s": ceiling,
            "achieved_cell_updates_per_s": achieved, "frac": achieved / ceiling,
            "issue_busy_measured": d.get("valu_issue_busy"),
            "source": "profiles/godunov_fp64_turb256_counters.json (SQ_INSTS_VALU per dispatch / cells)"}


def load_traffic(shape):
    """HBM bytes per launch of the FP64 kernel from the committed counter
    summary, when it was measured on the brick shape this run uses."""
    path = os.path.join(ROOT, "profiles", "godunov_fp64_turb256_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("brick_shape") != "%dx%dx%d" % shape:
        return None
    return d.get("hbm_bytes_per_launch")


def run_gpu(args, p, cells, mesh, precision, brick=None, kernel_steps=10):
    """Warm-up steps, then the K timed steps replayed as the captured hipGraph
    of a step pair (captured during the warm-up; --eager times them launch by
    launch), then a separate eager pass of kernel_steps steps whose Godunov
    launches are timed with HIP events on the library stream."""
    from ramses2_amd import UniformLevel, _lib
    from ramses2_amd.hydro import HipHydro
    if brick is not None:
        HipHydro.set_brick_shape(*brick)
    lev = UniformLevel(mesh, cells, p)
    graph = not args.eager
    if args.warmup:
        lev.run_device(max(args.warmup, 2) if graph else args.warmup, precision=precision, use_graph=graph)
    _lib.call("rh_device_sync")
    t0 = time.perf_counter()
    dt_last, t_tot, sums, _ = lev.run_device(args.steps, precision=precision, use_graph=graph)
    _lib.call("rh_device_sync")
    el = time.perf_counter() - t0
    _, _, _, kms = lev.run_device(kernel_steps, precision=precision, time_kernel=True)
    return lev, el, kms, dt_last, t_tot, sums


def precision_study(args, p, cells, mesh, bricks, nsteps=100):
    """The turbulence problem run nsteps steps in FP64 and in mixed precision
    from the same initial state, each with its own Courant dt: L1 of each
    conserved variable against FP64, and the drift of the total mass and
    energy over the run in each precision."""
    from ramses2_amd import UniformLevel
    from ramses2_amd.hydro import HipHydro
    out, states = {}, {}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        HipHydro.set_brick_shape(*bricks[prec])
        lev = UniformLevel(mesh, cells, p)
        c0 = lev.courant()
        _, t_tot, _, _ = lev.run_device(nsteps, precision=prec, use_graph=True)
        states[name] = lev.state()
        c1 = lev.courant()  # the same FP64 reduction as c0
        out[name] = {"t_end": t_tot, "mass_drift": float((c1["mass"] - c0["mass"]) / c0["mass"]),
                     "energy_drift": float((c1["ekin"] - c0["ekin"]) / c0["ekin"])}
    names = ["rho", "mom_x", "mom_y", "mom_z", "E"]
    u64, um = states["fp64"], states["mixed"]
    out["mixed"]["l1_rel_vs_fp64"] = {names[v]: float(np.abs(um[:, v] - u64[:, v]).sum() / np.abs(u64[:, v]).sum())
                                      for v in range(5)}
    out["steps"] = nsteps
    out["note"] = ("L1 = sum|mixed - fp64| / sum|fp64| over all cells after %d steps, each run with its own "
                   "Courant dt; drift = relative change of the total over the run (periodic box)" % nsteps)
    return out


def sedov_uniform(args):
    """Uniform 3D Sedov blast, 128^3 (levelmin = levelmax = 7) on the periodic
    unit box, gamma 1.4, ambient rho 1, p 1e-5, E = 1 in the central 2^3
    cells, MonCen, HLLC, Courant 0.5, sedov_steps steps in FP64 on the GPU
    timed after the warm-up steps (t_end is the time the timed steps advance);
    the CPU oracle on the same problem at all host threads and at one."""
    from ramses2_amd import HydroParams, UniformLevel, UniformMesh, _lib, ic
    from ramses2_amd.autotune import DEFAULT_SHAPE
    from ramses2_amd.hydro import HipHydro
    n = 128
    mesh = UniformMesh(3, 7)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    cells = ic.sedov_3d(n)
    ncells = mesh.noct * 8
    shape, shape_ms = DEFAULT_SHAPE, {}
    if not args.no_tune:
        from ramses2_amd.autotune import tune_uniform
        shape, shape_ms = tune_uniform(mesh, cells, p, precision=0, cache=None)
        print("bench: Sedov 128^3 brick shapes (Godunov ms): %s -> %s"
              % ({"%dx%dx%d" % k: round(v, 3) for k, v in shape_ms.items()}, shape), file=sys.stderr, flush=True)
    HipHydro.set_brick_shape(*shape)
    lev = UniformLevel(mesh, cells, p)
    lev.run_device(max(args.warmup, 2), use_graph=True)
    _lib.call("rh_device_sync")
    t0 = time.perf_counter()
    _, t_tot, _, _ = lev.run_device(args.sedov_steps, use_graph=True)
    _lib.call("rh_device_sync")
    el = time.perf_counter() - t0
    _, _, _, kms = lev.run_device(10, time_kernel=True)
    out = {"workload": "3D Sedov blast, uniform 128^3 (levelmin=levelmax=7), periodic unit box, gamma 1.4, "
                       "rho 1, p 1e-5, E=1 in the central 2^3 cells, MonCen, HLLC, Courant 0.5",
           "steps": args.sedov_steps, "warmup_steps": max(args.warmup, 2), "t_end": t_tot + 0.0,
           "gpu": {"cell_updates_per_s": ncells * args.sedov_steps / el, "ms_per_step": el / args.sedov_steps * 1e3,
                   "godunov_kernel_ms": kms,
                   "hbm_gbs_algorithmic": BYTES_PER_CELL * ncells / (kms * 1e-3) / 1e9 if kms else None,
                   "brick_shape": "%dx%dx%d" % shape,
                   "brick_shape_godunov_ms": {"%dx%dx%d" % k: v for k, v in shape_ms.items()} or None}}
    if not args.no_cpu:
        out["cpu"] = {}
        for threads in (len(os.sched_getaffinity(0)), 1):
            r = cpu_rate(mesh, cells, p, threads, args.cpu_seconds / 3)
            out["cpu"]["threads_%d" % threads] = r
        out["gpu_over_cpu_all_threads"] = out["gpu"]["cell_updates_per_s"] / out["cpu"][
            "threads_%d" % len(os.sched_getaffinity(0))]["cell_updates_per_s"]
    print("bench: Sedov 128^3 FP64 %.3f ms/step (%.3g cell-updates/s, kernel %.3f ms, %.0f GB/s algorithmic)"
          % (out["gpu"]["ms_per_step"], out["gpu"]["cell_updates_per_s"], kms or -1,
             out["gpu"]["hbm_gbs_algorithmic"] or -1), file=sys.stderr, flush=True)
    return out


def run_amr(args):
    """AMR Sedov (levels levelmin..levelmax refined on density/pressure
    gradients > 0.1, MonCen, HLLC, C 0.5): the batched step (every level in one
    work-list launch, hipGraph-replayed, brick shape from the autotuner) in FP64
    and mixed precision, beside the per-level drop-in call sequence."""
    from ramses2_amd.autotune import DEFAULT_SHAPE
    from ramses2_amd import HydroParams
    from ramses2_amd.amr_build import sedov_amr
    from ramses2_amd.autotune import tune_amr
    from ramses2_amd.hydro import AmrHierarchy, HipHydro

    lmin, lmax = (int(x) for x in args.amr_levels.split(","))
    t0 = time.perf_counter()
    am, u0 = sedov_amr(levelmin=lmin, levelmax=lmax, radius=0.3)
    t_build = time.perf_counter() - t0
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    cells = int(sum(am.noct[l] for l in am.levels) * 8)
    print("bench: AMR Sedov levels %d-%d, octs per level %s, %d cells (mesh built in %.1f s)"
          % (lmin, lmax, {l: am.noct[l] for l in am.levels}, cells, t_build), file=sys.stderr, flush=True)
    if args.amr_brick:
        bricks, times = {l: tuple(int(x) for x in args.amr_brick.split(",")) for l in am.levels}, {}
    elif args.no_tune:
        bricks, times = {l: DEFAULT_SHAPE for l in am.levels}, {}
    else:
        bricks, times = tune_amr(am, u0, p, precision=0, cache=None, key="sedov-amr")
    print("bench: AMR brick shapes (ms/step): %s -> %s" % ({k: round(v, 3) for k, v in times.items()}, bricks),
          file=sys.stderr, flush=True)
    # mixed precision has its own best shape (its kernel has other register and LDS limits)
    bricks_mixed, times_mixed = bricks, {}
    if not args.amr_brick and not args.no_tune:
        bricks_mixed, times_mixed = tune_amr(am, u0, p, precision=1, cache=None, key="sedov-amr")
        print("bench: AMR mixed brick shapes (ms/step): %s -> %s"
              % ({k: round(v, 3) for k, v in times_mixed.items()}, bricks_mixed), file=sys.stderr, flush=True)
    out = {"workload": "3D Sedov AMR, levels %d-%d, refinement on relative density/pressure gradient > 0.1 "
                       "(hydro_refine criterion on the analytic blast of radius 0.3), gamma 1.4, MonCen, HLLC, "
                       "Courant 0.5, periodic, nsubcycle=1" % (lmin, lmax),
           "octs_per_level": {str(l): int(am.noct[l]) for l in am.levels}, "cells": cells,
           "brick_shapes": {str(l): "%dx%dx%d" % b for l, b in bricks.items()},
           "brick_shape_ms_per_step": times or None,
           "brick_shapes_mixed": {str(l): "%dx%dx%d" % b for l, b in bricks_mixed.items()},
           "brick_shape_ms_per_step_mixed": times_mixed or None}
    states = {}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        h = AmrHierarchy(am, u0, p, bricks=bricks if prec == 0 else bricks_mixed)
        h.run_device(max(args.warmup, 2), precision=prec, use_graph=True)
        dt, tt, sums, ms = h.run_device(args.amr_steps, precision=prec, use_graph=True)
        states[name] = (h.state(), sums)
        if prec == 0:
            out["flags_after_run"] = amr_flags(h, am, p)
        out[name] = {"ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3), "steps": args.amr_steps}
        print("bench: AMR %s %.3f ms/step (%.3g cell-updates/s)" % (name, ms, cells / (ms * 1e-3)),
              file=sys.stderr, flush=True)
    out["mixed"]["speedup_vs_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    u64, s64 = states["fp64"]
    um, sm = states["mixed"]
    leaf = am.refined == 0
    l1 = {}
    for v, nm in ((0, "rho"), (4, "E")):
        a, b = u64[:, v][leaf], um[:, v][leaf]
        l1[nm] = float(np.abs(a - b).sum() / np.abs(a).sum())
    out["mixed"]["l1_rel_vs_fp64_leaf"] = l1
    out["mixed"]["mass_drift_vs_fp64"] = float((sm[0] - s64[0]) / s64[0])
    # the drop-in sequence: per level courant (synchronous), set_unew, godunov, set_uold, upload
    h = AmrHierarchy(am, u0, p, bricks=bricks)
    h.step()
    from ramses2_amd import _lib
    _lib.call("rh_device_sync")
    t0 = time.perf_counter()
    nd = max(3, args.amr_steps // 4)
    for _ in range(nd):
        h.step()
    _lib.call("rh_device_sync")
    ms_d = (time.perf_counter() - t0) / nd * 1e3
    out["dropin_per_level_fp64"] = {"ms_per_step": ms_d, "cell_updates_per_s": cells / (ms_d * 1e-3),
                                    "steps": nd,
                                    "note": "reference call order per level through the drop-in C-ABI"}
    out["fp64"]["speedup_vs_dropin"] = ms_d / out["fp64"]["ms_per_step"]
    print("bench: AMR drop-in per level %.3f ms/step; batched FP64 x%.2f, mixed x%.2f vs FP64"
          % (ms_d, out["fp64"]["speedup_vs_dropin"], out["mixed"]["speedup_vs_fp64"]), file=sys.stderr, flush=True)
    HipHydro.set_brick_shape(*DEFAULT_SHAPE)
    return out


def amr_flags(h, am, p, err=0.1):
    """The refinement criterion (rh_hydro_flag: relative density and pressure
    gradients > err) on the device state after the timed steps, per level:
    cells flagged, flagged cells without children (a refine step would split
    them) and refined cells not flagged (candidates for derefinement)."""
    t0 = time.perf_counter()
    out = {}
    for l in am.levels:
        f, n = h.hh.hydro_flag(am.head[l], am.noct[l], p, err, err)
        ref = am.refined[am.head[l] - 1:am.head[l] - 1 + am.noct[l]] != 0
        out[str(l)] = {"flagged": int(n), "flagged_leaf": int(((f != 0) & ~ref).sum()),
                       "refined_unflagged": int((ref & (f == 0)).sum())}
    ms = (time.perf_counter() - t0) * 1e3
    print("bench: AMR flags from the evolved state (err_grad_d = err_grad_p = %g): %s (%.1f ms)"
          % (err, out, ms), file=sys.stderr, flush=True)
    return {"err_grad_d": err, "err_grad_p": err, "per_level": out, "ms_all_levels": ms}


def run_serving(args):
    """Ensemble serving: capacity of the batched FP64 step measured on a closed
    batch of the stream's first jobs (brick shape: the fastest tick of that
    batch), then the Poisson stream of all jobs at serve_load of that
    capacity, served in FP64 and in mixed precision, each with the brick
    shape whose engine served the stream's first jobs with the lowest mean
    latency (tune_serving)."""
    from ramses2_amd.autotune import DEFAULT_SHAPE
    from ramses2_amd import _lib
    from ramses2_amd.autotune import tune_serving
    from ramses2_amd.serving import EnsembleServer, HipEngine, latency_summary, make_stream, mean_job_cells

    base = make_stream(args.serve_jobs, rate=1.0, seed=args.seed, nsteps=args.serve_steps)
    budget = int(args.serve_budget_gb * (1 << 30))
    # capacity: serve_slots jobs of the stream resident, 20 batched steps timed
    # (admission timed apart: it is host work between ticks)
    cal = base[:args.serve_slots]
    # brick shape for the serving mix: batched FP64 ticks of the calibration jobs per shape
    shape_ms = {}
    if args.serve_brick:
        brick = tuple(int(x) for x in args.serve_brick.split(","))
    elif not args.no_tune:
        from ramses2_amd.autotune import compiled_shapes
        for sh in compiled_shapes():
            e = HipEngine(precision=0, brick=sh)
            hs = [e.open(s) for s in cal]
            e.step(hs, 2)
            e.sync()
            t0 = time.perf_counter()
            e.step(hs, 10)
            e.sync()
            shape_ms[sh] = (time.perf_counter() - t0) / 10 * 1e3
            for h in hs:
                e.close(h)
            e.close_all()
        brick = min(shape_ms, key=shape_ms.get)
        print("bench: serving brick shapes (ms per tick of %d jobs): %s -> %s" % (
            len(cal), {"%dx%dx%d" % k: round(v, 3) for k, v in shape_ms.items()}, brick), file=sys.stderr, flush=True)
    else:
        brick = DEFAULT_SHAPE
    eng = HipEngine(precision=0, brick=brick)
    # one template context per grid size of the stream (the server's start-up)
    t0 = time.perf_counter()
    for n in sorted({s.ncell for s in base}):
        eng.warm(n)
    eng.sync()
    t_warm = time.perf_counter() - t0
    t0 = time.perf_counter()
    hs = [eng.open(s) for s in cal]
    eng.sync()
    t_open = (time.perf_counter() - t0) / len(cal)
    eng.step(hs, 2)
    eng.sync()
    t0 = time.perf_counter()
    ncal = 20
    for _ in range(ncal):
        eng.step(hs)
    eng.sync()
    el = time.perf_counter() - t0
    t0 = time.perf_counter()
    for h in hs:
        eng.close(h)
    t_close = (time.perf_counter() - t0) / len(cal)
    capacity = sum(s.cells for s in cal) * ncal / el
    # the same calibration batch stepped in mixed precision
    engm = HipEngine(precision=1, brick=brick)
    hs = [engm.open(s) for s in cal]
    engm.step(hs, 2)
    engm.sync()
    t0 = time.perf_counter()
    for _ in range(ncal):
        engm.step(hs)
    engm.sync()
    capacity_mixed = sum(s.cells for s in cal) * ncal / (time.perf_counter() - t0)
    for h in hs:
        engm.close(h)
    engm.close_all()
    mean_cells = mean_job_cells(base)
    rate = args.serve_rate or args.serve_load * capacity / (mean_cells * args.serve_steps)
    for s in base:
        s.arrival = s.arrival / rate
    print("bench: serving capacity %.3g cell-updates/s FP64, %.3g mixed (%d jobs resident, batched steps); "
          "admission %.2f ms, release %.2f ms per job; rate %.2f jobs/s for load %.2f of FP64"
          % (capacity, capacity_mixed, len(cal), t_open * 1e3, t_close * 1e3, rate,
             rate * mean_cells * args.serve_steps / capacity),
          file=sys.stderr, flush=True)
    out = {"jobs": len(base), "arrival_rate_jobs_per_s": rate,
           "offered_load": rate * mean_cells * args.serve_steps / capacity,
           "capacity_cell_updates_per_s": capacity, "capacity_mixed_cell_updates_per_s": capacity_mixed,
           "capacity_note": "%d jobs of the stream resident, %d batched steps timed in each precision; both "
           "precisions serve the stream at the FP64-derived rate" % (len(cal), ncal), "admission_ms_per_job": t_open * 1e3,
           "release_ms_per_job": t_close * 1e3, "template_build_s": t_warm,
           "template_bytes": eng.template_bytes(), "mean_job_cells": mean_cells, "steps_per_job": args.serve_steps,
           "max_slots": args.serve_slots, "pipeline_depth_ticks": args.serve_depth, "hbm_budget_bytes": budget, "capacity_brick_shape": "%dx%dx%d" % brick,
           "brick_shape_ms_per_tick": {"%dx%dx%d" % k: v for k, v in shape_ms.items()} or None,
           "sizes": "N^3, N log-uniform in [32,128], even; 50% blast, 50% Riemann; gamma 1.4, MonCen, HLLC, C 0.5"}
    eng.close_all()
    for prec, name in ((0, "fp64"), (1, "mixed")):
        # the shape for the stream itself: the shapes within 1.3x of the best
        # capacity tick serve the stream's first jobs at its arrival times
        shape_p, trial = brick, {}
        if shape_ms and args.serve_trial_jobs > 0:
            best_ms = min(shape_ms.values())
            cands = [sh for sh, ms in sorted(shape_ms.items(), key=lambda x: x[1]) if ms <= 1.3 * best_ms]
            shape_p, trial = tune_serving(base[:args.serve_trial_jobs], prec, cands, max_slots=args.serve_slots,
                                          hbm_budget=budget, depth=args.serve_depth)
            print("bench: serving %s brick shapes (mean latency of the first %d jobs, s): %s -> %s"
                  % (name, args.serve_trial_jobs, {"%dx%dx%d" % k: round(v, 4) for k, v in trial.items()}, shape_p),
                  file=sys.stderr, flush=True)
        engp = HipEngine(precision=prec, brick=shape_p)
        for n in sorted({s.ncell for s in base}):
            engp.warm(n)
        srv = EnsembleServer(engp, max_slots=args.serve_slots, hbm_budget=budget, depth=args.serve_depth)
        res = srv.run(base)
        summ = latency_summary(res, base)
        tr = np.array(srv.trace)
        summ.update({"ticks": srv.ticks, "mean_resident_jobs": float(tr[:, 1].mean()) if len(tr) else 0.0,
                     "max_bytes_in_use": int(tr[:, 2].max()) if len(tr) else 0,
                     "brick_shape": "%dx%dx%d" % shape_p,
                     "brick_shape_trial_mean_latency_s": {"%dx%dx%d" % k: v for k, v in trial.items()} or None})
        out[name] = summ
        print("bench: serving %s p50 %.3f s p99 %.3f s, %.3g cell-updates/s over %.1f s"
              % (name, summ["p50_s"], summ["p99_s"], summ["throughput_cell_updates_per_s"], summ["span_s"]),
              file=sys.stderr, flush=True)
        engp.close_all()
    _lib.call("rh_ctx_select", 0)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: the hydro path runs on one GPU (one MI355X); multi-GPU sharding is a design "
              "only, so --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    from ramses2_amd import HydroParams, UniformMesh, ic

    n = args.ncell
    level = int(np.log2(n))
    assert 2 ** level == n, "ncell must be a power of two"
    p = HydroParams(gamma=5.0 / 3.0, courant_factor=0.5, slope_type=2, riemann="hllc")
    mesh = UniformMesh(3, level)
    cells = ic.turbulence_3d(n, seed=args.seed)
    ncells = mesh.noct * 8
    print("bench: turbulence %d^3 (%d cells), warmup %d, steps %d" % (n, ncells, args.warmup, args.steps),
          file=sys.stderr, flush=True)

    if args.copy_probe_mb > 0:
        import ctypes
        from ramses2_amd import _lib
        _lib.call("rh_init")
        ms = ctypes.c_float()
        nbytes = args.copy_probe_mb << 20
        _lib.call("rh_copy_probe", ctypes.c_int64(nbytes), ctypes.byref(ms))
        print("bench: copy probe %d B read + %d B written in %.3f ms (%.0f GB/s)"
              % (nbytes, nbytes, ms.value, 2 * nbytes / (ms.value * 1e-3) / 1e9), file=sys.stderr, flush=True)

    from ramses2_amd.autotune import DEFAULT_SHAPE
    bricks = {0: DEFAULT_SHAPE, 1: DEFAULT_SHAPE}
    fixed = {}
    for prec, opt in ((0, args.fp64_brick), (1, args.mixed_brick)):
        if opt:
            fixed[prec] = bricks[prec] = tuple(int(x) for x in opt.split(","))
    tuning = {}
    if not args.no_tune:
        from ramses2_amd.autotune import tune_uniform
        for prec in [q for q in ((0, 1) if not args.no_mixed else (0,)) if q not in fixed]:
            best, times = tune_uniform(mesh, cells, p, precision=prec, steps=6, cache=None)
            bricks[prec] = best
            tuning["fp64" if prec == 0 else "mixed"] = {"best": "%dx%dx%d" % best, "ms_per_step": {
                "%dx%dx%d" % k: v for k, v in times.items()}}
            print("bench: brick shapes %s (Godunov ms): %s -> %s" % (
                "fp64" if prec == 0 else "mixed", {"%dx%dx%d" % k: round(v, 3) for k, v in times.items()}, best),
                file=sys.stderr, flush=True)

    lev64, el64, kms64, dt64, t64, sums64 = run_gpu(args, p, cells, mesh, precision=0, brick=bricks[0])
    value = ncells * args.steps / el64
    ms_per_step = el64 / args.steps * 1e3
    print("bench: fp64 %.3f ms/step, godunov kernel %.3f ms" % (ms_per_step, kms64 or -1), file=sys.stderr,
          flush=True)

    mixed = None
    if not args.no_mixed:
        levm, elm, kmsm, dtm, tm, sumsm = run_gpu(args, p, cells, mesh, precision=1, brick=bricks[1])
        mixed = {"value": ncells * args.steps / elm, "ms_per_step": elm / args.steps * 1e3,
                 "speedup_vs_fp64": el64 / elm, "godunov_kernel_ms": kmsm,
                 "kernel_speedup_vs_fp64": (kms64 / kmsm) if (kms64 and kmsm) else None}
        print("bench: mixed %.3f ms/step (x%.2f), godunov kernel %.3f ms"
              % (mixed["ms_per_step"], mixed["speedup_vs_fp64"], kmsm or -1), file=sys.stderr, flush=True)
        study = precision_study(args, p, cells, mesh, bricks, nsteps=args.study_steps)
        mixed["precision_study"] = study
        print("bench: precision study %d steps: L1 mixed vs FP64 %s; drift mass %.2g/%.2g energy %.2g/%.2g "
              "(FP64/mixed)" % (study["steps"], {k: "%.2g" % v for k, v in study["mixed"]["l1_rel_vs_fp64"].items()},
                                study["fp64"]["mass_drift"], study["mixed"]["mass_drift"],
                                study["fp64"]["energy_drift"], study["mixed"]["energy_drift"]),
              file=sys.stderr, flush=True)

    roof = None
    if kms64:
        achieved = BYTES_PER_CELL * ncells / (kms64 * 1e-3) / 1e9
        traffic = load_traffic(bricks[0])
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": "k_godunov3d<%d,%d,%d,double> (FP64)" % bricks[0], "kernel_ms": kms64,
                "bytes_per_launch": BYTES_PER_CELL * ncells}
        valu = valu_ceiling(bricks[0], ncells, kms64)
        if valu:
            roof["valu"] = valu
            roof["limiter"] = "valu-issue" if valu["frac"] > roof["frac"] else "hbm"

    sedov = None
    if not args.no_sedov:
        sedov = sedov_uniform(args)

    amr = None
    if not args.no_amr:
        amr = run_amr(args)

    serving = None
    if not args.no_serve:
        serving = run_serving(args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args, p)
        print("bench: cpu %.3g cell-updates/s on %d threads" % (cpu["value"], cpu["cores"]), file=sys.stderr,
              flush=True)

    out = {
        "metric": "cell-updates/s 3D Godunov (FP64 vs mixed) at 256^3; p50/p99 ensemble job latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D supersonic decaying turbulence, uniform %d^3, gamma=5/3, MonCen, HLLC, "
                               "Courant 0.5, periodic; full level step (courant_fine + godunov_fine)" % n,
                   "cells": ncells, "octs": mesh.noct, "parallelism": "single GPU",
                   "brick_shape": {"fp64": "%dx%dx%d" % bricks[0], "mixed": "%dx%dx%d" % bricks[1]},
                   "autotune": tuning or None},
        "mixed": mixed,
        "roofline": roof,
        "cpu_baseline": cpu,
        "sedov_128": sedov,
        "amr": amr,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
