"""Brick-shape autotuner.

The Godunov kernels process octs in bricks of BX x BY x BZ octs (one
workgroup per brick, (2BX+2)(2BY+2) threads marching in z).  The best shape
depends on the mesh (uniform box, sparse AMR shells, small serving jobs), the
precision and the device, so it is measured: each compiled shape is timed on
the actual mesh with HIP events, and the fastest is kept.  Results are cached
per (workload key, precision) in a JSON file next to the library unless a
path is given.
"""
import ctypes
import json
import os

import numpy as np

from . import _lib
from ._lib import call

# the library's shape until one is set (rh_state.h)
DEFAULT_SHAPE = (8, 6, 8)
SHAPES = [(7, 7, 8), (7, 7, 16), (8, 6, 8), (8, 6, 16), (4, 4, 4), (4, 4, 8), (15, 7, 8), (15, 7, 16), (7, 7, 4),
          (3, 3, 8), (3, 3, 4)]
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "autotune.json")


def compiled_shapes():
    """The shapes of SHAPES that the loaded library has kernels for."""
    out = []
    for s in SHAPES:
        v = [ctypes.c_int32() for _ in range(3)]
        call("rh_get_brick_shape", *[ctypes.byref(x) for x in v])
        old = tuple(x.value for x in v)
        ok = _lib.load().rh_set_brick_shape(*s) == 0
        call("rh_set_brick_shape", *old)
        if ok:
            out.append(s)
    return out


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, d):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(key, precision, run, shapes=None, cache=DEFAULT_CACHE, refresh=False):
    """Time run(shape) -> ms for each shape (run builds the mesh tables with
    that shape and returns the mean step time); return (best_shape, {shape: ms}).
    A cached choice for (key, precision) is returned without timing unless
    refresh is set."""
    ck = "%s|prec%d" % (key, precision)
    db = _load(cache) if cache else {}
    if not refresh and ck in db:
        return tuple(db[ck]["best"]), {tuple(map(int, k.split("x"))): v for k, v in db[ck]["ms"].items()}
    shapes = shapes or compiled_shapes()
    times = {}
    for s in shapes:
        times[s] = float(run(s))
    best = min(times, key=times.get)
    if cache:
        db = _load(cache)
        db[ck] = {"best": list(best), "ms": {"%dx%dx%d" % s: t for s, t in times.items()}}
        _store(cache, db)
    return best, times


def tune_uniform(mesh, cells, params, precision=0, steps=6, shapes=None, cache=DEFAULT_CACHE, refresh=False):
    """Best brick shape for device-resident steps of a uniform level."""
    from .hydro import HipHydro, UniformLevel

    def run(shape):
        HipHydro.set_brick_shape(*shape)
        lev = UniformLevel(mesh, cells, params)
        lev.run_device(1, precision=precision)
        _, _, _, ms = lev.run_device(steps, precision=precision, time_kernel=True)
        return ms

    key = "uniform%d^3|slope%d|%s" % (mesh.n * 2, params.slope_type, params.riemann)
    best, times = tune(key, precision, run, shapes, cache, refresh)
    HipHydro.set_brick_shape(*best)
    return best, times


def tune_amr(amesh, uold, params, precision=0, steps=4, shapes=None, cache=DEFAULT_CACHE, refresh=False,
             interpol_var=0, interpol_type=1, key=None):
    """Brick shapes for batched steps of an AMR mesh: every compiled shape on
    all levels, and the per-level assignment of level_shapes, each timed on
    the mesh; returns ({level: shape}, {candidate: ms per step})."""
    from .hydro import AmrHierarchy, HipHydro

    per_level = level_shapes(amesh, shapes)

    def run(cand):
        bricks = per_level if cand == "per-level" else {l: cand for l in amesh.levels}
        h = AmrHierarchy(amesh, uold, params, interpol_var=interpol_var, interpol_type=interpol_type, bricks=bricks)
        h.run_device(2, precision=precision, use_graph=True)
        return h.run_device(steps, precision=precision, use_graph=True)[3]

    key = key or "amr%d-%d|octs%d|slope%d|%s" % (amesh.levelmin, amesh.levelmax, amesh.ngrid, params.slope_type,
                                                   params.riemann)
    cands = list(shapes or compiled_shapes()) + ["per-level"]
    ck = "%s|prec%d" % (key, precision)
    db = _load(cache) if cache else {}
    if not refresh and ck in db:
        times = db[ck]["ms"]
    else:
        times = {("%dx%dx%d" % c if c != "per-level" else c): float(run(c)) for c in cands}
        if cache:
            db = _load(cache)
            db[ck] = {"ms": times}
            _store(cache, db)
    best = min(times, key=times.get)
    bricks = per_level if best == "per-level" else {l: tuple(int(x) for x in best.split("x")) for l in amesh.levels}
    HipHydro.set_brick_shape(*bricks[amesh.levelmin])
    return bricks, times


def tune_serving(specs, precision, shapes, max_slots=64, hbm_budget=96 << 30, params=None, depth=2):
    """Brick shape for serving a request stream at its arrival times: each
    shape serves the same stream prefix (specs, arrival times in seconds) with
    its own engine, and the shape with the lowest mean latency is kept.  A
    tick's length depends on how well the resident jobs' bricks fill the GPU,
    and the jobs resident under load are fewer than a capacity batch, so the
    fastest shape for a full batch need not be the fastest under the stream.
    Returns (best_shape, {shape: mean latency in s})."""
    from .serving import EnsembleServer, HipEngine

    lat = {}
    for sh in shapes:
        eng = HipEngine(params, precision=precision, brick=sh)
        try:
            for n in sorted({s.ncell for s in specs}):
                eng.warm(n)
            res = EnsembleServer(eng, max_slots=max_slots, hbm_budget=hbm_budget, depth=depth).run(specs)
            lat[tuple(sh)] = float(np.mean([r.latency for r in res.values()]))
        finally:
            eng.close_all()
    best = min(lat, key=lat.get)
    return best, lat


def brick_offset(keys, shape, box_min=(0, 0, 0), box_max=None):
    """The lattice offset rh_build_nbor gives a level's bricks of this shape
    (origins at box_min + offset + multiples of the shape) and its modelled
    work, from the library's own choice (rh_brick_lattice, host only).  With
    no box_max the level is taken as sparse in an unbounded box."""
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.int32).reshape(-1, 3))
    bmin = np.ascontiguousarray(np.asarray(box_min, dtype=np.int32)[:3])
    if box_max is None:
        box_max = np.full(3, 1 << 30, dtype=np.int32)
    bmax = np.ascontiguousarray(np.asarray(box_max, dtype=np.int32)[:3])
    off = np.zeros(3, dtype=np.int32)
    work = ctypes.c_double()
    ip = ctypes.POINTER(ctypes.c_int32)
    call("rh_brick_lattice", k.ctypes.data_as(ip), len(k), bmin.ctypes.data_as(ip), bmax.ctypes.data_as(ip),
         int(shape[0]), int(shape[1]), int(shape[2]), off.ctypes.data_as(ip), ctypes.byref(work))
    return tuple(int(x) for x in off), work.value


def level_work(keys, shape, box_min=(0, 0, 0), box_max=None):
    """Modelled work of one level's bricks on the lattice rh_build_nbor picks
    (brick_offset): a brick marches the planes of its occupied z layers plus
    two halo planes each side; per plane the P phase converts (2BX+4)(2BY+4)
    cells and the T/F/U phases run (2BX+2)(2BY+2) columns, a P cell costing
    about a third of a T/F/U column-plane (0.35)."""
    return brick_offset(keys, shape, box_min, box_max)[1]


def count_bricks(keys, shape, box_min=(0, 0, 0)):
    """Bricks holding at least one of the octs (oct keys, (n,3))."""
    k = (np.asarray(keys, dtype=np.int64) - np.asarray(box_min)) // np.asarray(shape)
    return len(np.unique(k[:, 0] | (k[:, 1] << 21) | (k[:, 2] << 42)))


def level_shapes(amesh, shapes=None):
    """Per level, the shape with the least modelled work (level_work)."""
    shapes = shapes or [s for s in SHAPES if s != (1, 1, 1)]
    _, _, box_min, box_max = amesh.level_params()
    out = {}
    for l in amesh.levels:
        i0 = amesh.head[l] - 1
        k = amesh.ckey[i0:i0 + amesh.noct[l]]
        out[l] = min(shapes, key=lambda s: level_work(k, s, box_min[l - 1], box_max[l - 1]))
    return out


def fill_efficiency(noct, n_bricks, shape):
    """Fraction of brick slots holding an oct of the level."""
    return noct / float(n_bricks * int(np.prod(shape)))
