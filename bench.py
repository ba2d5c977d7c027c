#!/usr/bin/env python3
"""Throughput of the fine-level Godunov hydro step on one MI355X.

Workload: uniform 256^3 supersonic decaying turbulence (gamma=5/3, rho=1, p=1,
solenoidal velocity with P(k)~k^-2 on 2<=k<=8 at rms Mach 5, seeded), MonCen
slopes, HLLC, Courant 0.5, periodic.  A step is one full level step: the
Courant dt reduction (courant_fine) and the Godunov update (godunov_fine with
set_unew/set_uold folded in), all on the device.  The state is resident in HBM
before the timed region starts.

Prints one JSON line: value = FP64 cell-updates/s of the whole step; the mixed
precision (FP32 reconstruction/fluxes, FP64 accumulation) rate and its L1
error against the FP64 run sit beside it; roofline = the Godunov kernel's
algorithmic HBM bytes (80 B per cell-update) over its mean launch time measured
with HIP events inside the timed steps; cpu_baseline = the CPU oracle on the
host cores on a bounded sample.
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
BYTES_PER_CELL = 80.0  # read uold 5x8 B + write unew 5x8 B


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ncell", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-sample", type=int, default=128, help="cells per dimension of the CPU sample")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay steps as a hipGraph (no per-kernel events)")
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


def cpu_baseline(args, p):
    """The CPU oracle (C restatement of godunov_fine/courant_fine, OpenMP over
    octs) on a bounded sample: the same problem at cpu_sample^3 cells."""
    from oracle import oracle as O
    from ramses2_amd import UniformMesh, ic
    from tests.helpers import oracle_params, oracle_uniform

    n = args.cpu_sample
    level = int(np.log2(n))
    mesh = UniformMesh(3, level)
    cells = ic.turbulence_3d(n, seed=args.seed)
    om = oracle_uniform(mesh, mesh.from_cells(cells))
    po = oracle_params(p)
    octmap = (mesh.octmap + 1).astype(np.int32)
    threads = len(os.sched_getaffinity(0))
    O.lib()

    def one_step():
        c = om.courant_fine(level, po, threads)
        om.set_unew(level)
        om.godunov_uniform_omp(level, c[4], po, octmap, threads)
        om.set_uold(level)

    t0 = time.perf_counter()
    one_step()
    t1 = time.perf_counter()
    nsteps = max(1, min(200, int(args.cpu_seconds / max(t1 - t0, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(nsteps):
        one_step()
    el = time.perf_counter() - t0
    rate = nsteps * mesh.noct * 8 / el
    return {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle (oracle/ramses_oracle.c, OpenMP over octs) on the same turbulence problem "
                      "at %d^3 cells, %d full steps (courant_fine + godunov_fine), %.1f s; host CPU: %s"
                      % (n, nsteps, el, cpu_model())}


def load_traffic():
    path = os.path.join(ROOT, "profiles", "godunov_fp64_turb256_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def run_gpu(args, p, cells, mesh, precision):
    from ramses2_amd import UniformLevel, _lib
    lev = UniformLevel(mesh, cells, p)
    if args.warmup:
        lev.run_device(args.warmup, precision=precision)
    _lib.call("rh_device_sync")
    t0 = time.perf_counter()
    dt_last, t_tot, sums, kms = lev.run_device(args.steps, precision=precision, use_graph=args.graph,
                                                time_kernel=not args.graph)
    _lib.call("rh_device_sync")
    el = time.perf_counter() - t0
    return lev, el, kms, dt_last, t_tot, sums


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (each GPU box has one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
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

    lev64, el64, kms64, dt64, t64, sums64 = run_gpu(args, p, cells, mesh, precision=0)
    u64 = lev64.state()
    value = ncells * args.steps / el64
    ms_per_step = el64 / args.steps * 1e3
    print("bench: fp64 %.3f ms/step, godunov kernel %.3f ms" % (ms_per_step, kms64 or -1), file=sys.stderr,
          flush=True)

    mixed = None
    if not args.no_mixed:
        levm, elm, kmsm, dtm, tm, sumsm = run_gpu(args, p, cells, mesh, precision=1)
        um = levm.state()
        l1 = {}
        names = ["rho", "mom_x", "mom_y", "mom_z", "E"]
        for v in range(5):
            l1[names[v]] = float(np.abs(um[:, v] - u64[:, v]).sum() / np.abs(u64[:, v]).sum())
        mixed = {"value": ncells * args.steps / elm, "ms_per_step": elm / args.steps * 1e3,
                 "speedup_vs_fp64": el64 / elm, "godunov_kernel_ms": kmsm,
                 "l1_rel_vs_fp64": l1,
                 "l1_note": "after %d steps, each run with its own Courant dt" % (args.warmup + args.steps),
                 "mass_drift": float((sumsm[0] - sums64[0]) / sums64[0])}
        print("bench: mixed %.3f ms/step (x%.2f)" % (mixed["ms_per_step"], mixed["speedup_vs_fp64"]),
              file=sys.stderr, flush=True)

    roof = None
    if kms64:
        achieved = BYTES_PER_CELL * ncells / (kms64 * 1e-3) / 1e9
        traffic = load_traffic()
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "kernel": "k_godunov3d<7,7,8,double> (FP64)", "kernel_ms": kms64,
                "bytes_per_launch": BYTES_PER_CELL * ncells}

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
                   "cells": ncells, "octs": mesh.noct, "parallelism": "single GPU"},
        "mixed": mixed,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
