"""Ensemble serving: a request loop over independent 3D hydro jobs.

Jobs arrive as a Poisson stream.  Each is a seeded random blast or a planar
two-state Riemann problem on a periodic N^3 grid (N even, log-uniform in
[nmin, nmax]), run for a fixed number of Courant-limited steps with gamma=1.4,
MonCen slopes and HLLC.  A job's state stays resident in HBM from admission to
completion, in its own library context (ramses2_hip_ext.h, rh_ctx_*).

Continuous batching: every tick advances all admitted jobs by one time step
with shared launches (rh_batch_step: one work-list Godunov kernel over the
bricks of every job, one per-job Courant reduction).  The host keeps one
tick in flight: while the device runs tick k, finished jobs of tick k-1 leave
(their statuses copied by the stream at a fence, rh_fence_record) and queued
jobs are admitted in arrival order while a slot is free and their device
footprint fits the HBM budget (rh_estimate_uniform_bytes).  Work of a job does
not depend on which other jobs share its launches, so a job's result is the
same alone and in the mix.

The loop is written against a small engine interface (open / step / fence /
wait / close) so the scheduling can be tested without a GPU; HipEngine is the
device engine.
"""
import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import call, ptr
from .hydro import HydroParams
from .mesh import NVAR, UniformMesh

KIND_BLAST, KIND_RIEMANN = 1, 2


@dataclass
class JobSpec:
    """One request: a seeded initial condition on an ncell^3 periodic box."""

    job_id: int
    arrival: float            # seconds after the start of the stream
    ncell: int                # cells per dimension (even)
    kind: int                 # KIND_BLAST or KIND_RIEMANN
    ic: tuple                 # the five device IC parameters of rh_init_ic
    nsteps: int = 100
    info: dict = field(default_factory=dict)

    @property
    def cells(self):
        return self.ncell ** 3


def job_ic(job_id, ncell, seed, gamma=1.4):
    """Random initial condition of one job from its own generator
    (seed + job_id).  Blast: E in U[0.1, 10], ambient p log-uniform in
    [1e-6, 1e-2], the energy deposited in the central 2^3 cells as pressure
    (gamma-1) E / V (as ic.sedov_3d).  Riemann: rho and p of each side
    log-uniform in [0.1, 10], interface normal to a random axis at the box
    centre (as ic.riemann_3d)."""
    rng = np.random.default_rng(seed + job_id)
    dx = 1.0 / ncell
    if rng.random() < 0.5:
        e = float(rng.uniform(0.1, 10.0))
        p0 = float(10 ** rng.uniform(-6, -2))
        pb = (gamma - 1.0) * e / (2 ** 3 * dx ** 3)
        lo = ncell // 2 - 1
        return KIND_BLAST, (1.0, p0, pb, float(lo), 0.0), {"kind": "blast", "E": e, "p0": p0}
    rl, rr = 10 ** rng.uniform(-1, 1, 2)
    pl, pr = 10 ** rng.uniform(-1, 1, 2)
    axis = int(rng.integers(0, 3))
    return (KIND_RIEMANN, (float(rl), float(pl), float(rr), float(pr), float(axis)),
            {"kind": "riemann", "axis": axis, "rho": (float(rl), float(rr)), "p": (float(pl), float(pr))})


def ic_cells(spec: JobSpec, gamma=1.4):
    """The job's initial conservative cells (5, N, N, N) on the host, from the
    same parameters as the device initial condition (tests compare the two)."""
    from .mesh import prim_to_cons
    n = spec.ncell
    a = spec.ic
    z = np.zeros((n, n, n))
    if spec.kind == KIND_BLAST:
        rho = np.full((n, n, n), a[0])
        p = np.full((n, n, n), a[1])
        lo = int(a[3])
        p[lo:lo + 2, lo:lo + 2, lo:lo + 2] = a[2]
    else:
        idx = np.arange(n)
        left = (idx + 0.5) / n < 0.5
        shape = [1, 1, 1]
        shape[2 - int(a[4])] = n
        L = np.broadcast_to(left.reshape(shape), (n, n, n))
        rho = np.where(L, a[0], a[2])
        p = np.where(L, a[1], a[3])
    return prim_to_cons(rho, (z, z, z), p, gamma)


def make_stream(njobs, rate, seed=20261202, nmin=32, nmax=128, nsteps=100, gamma=1.4):
    """njobs requests with exponential inter-arrival times of mean 1/rate
    (a Poisson stream), sizes N log-uniform in [nmin, nmax] rounded to an even
    number, half blasts and half Riemann problems."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, njobs) if rate > 0 else np.zeros(njobs)
    arrivals = np.cumsum(gaps) - gaps[0]
    logs = rng.uniform(np.log(nmin), np.log(nmax), njobs)
    specs = []
    for j in range(njobs):
        n = int(2 * round(np.exp(logs[j]) / 2))
        n = min(max(n, nmin), nmax)
        kind, a, info = job_ic(j, n, seed, gamma)
        specs.append(JobSpec(j, float(arrivals[j]), n, kind, a, nsteps, info))
    return specs


def mean_job_cells(specs):
    return float(np.mean([s.cells for s in specs]))


# ---------------------------------------------------------------------------
# Device engine
# ---------------------------------------------------------------------------
class HipEngine:
    """Jobs as library contexts on the GPU."""

    def __init__(self, params: HydroParams = None, precision=0, brick=None):
        self.p = params or HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
        self.precision = precision
        self._mesh = {}
        self._templates = {}
        self._tbytes = None
        call("rh_init")
        if brick is None:
            v = [ctypes.c_int32() for _ in range(3)]
            call("rh_get_brick_shape", *[ctypes.byref(x) for x in v])
            brick = tuple(x.value for x in v)
        self.brick = tuple(int(b) for b in brick)
        # context 0 holds the shape the footprint estimate uses
        call("rh_set_brick_shape", *self.brick)

    def _mesh_for(self, ncell):
        m = self._mesh.get(ncell)
        if m is None:
            mesh = UniformMesh.for_cells(3, ncell)
            lp = mesh.level_params()
            m = (mesh, mesh.grid_records(), [np.ascontiguousarray(x) for x in lp])
            self._mesh[ncell] = m
        return m

    def footprint(self, spec: JobSpec):
        """Device bytes the job will hold (the admission estimate)."""
        mesh = self._mesh_for(spec.ncell)[0]
        b = ctypes.c_int64()
        call("rh_estimate_uniform_bytes", mesh.n, 0, 2 * mesh.noct, ctypes.byref(b))
        return b.value

    def _build(self, ncell):
        """A context holding the mesh of an ncell^3 periodic box: level
        parameters, oct records, device hash, neighbour and brick tables."""
        mesh, grid, (ckey_max, key_off, box_min, box_max) = self._mesh_for(ncell)
        cid = ctypes.c_int32()
        call("rh_ctx_create", ctypes.byref(cid))
        cid = cid.value
        call("rh_ctx_select", cid)
        try:
            call("rh_init")
            call("rh_set_brick_shape", *self.brick)
            call("rh_alloc_amr", mesh.noct, 0, NVAR, 8, 2 * mesh.noct)
            per = np.ones(3, dtype=np.int32)
            call("rh_upload_level_params", ptr(ckey_max, _lib._ip), ptr(key_off, _lib._lp), ptr(box_min, _lib._ip),
                 ptr(box_max, _lib._ip), ptr(per, _lib._ip), len(ckey_max) - 1)
            call("rh_set_grid_device", None, None, None, grid.ctypes.data_as(ctypes.c_void_p), mesh.noct, NVAR, 8)
            call("rh_insert_hash_all", 1, mesh.noct, 2 * mesh.noct)
            l = mesh.level - 1
            bmin = np.ascontiguousarray(box_min[l])
            bmax = np.ascontiguousarray(box_max[l])
            call("rh_build_nbor", 1, mesh.noct, 2 * mesh.noct, int(ckey_max[l]), int(key_off[l]),
                 ptr(bmin, _lib._ip), ptr(bmax, _lib._ip), ptr(per, _lib._ip))
        except Exception:
            call("rh_ctx_destroy", cid)
            raise
        finally:
            call("rh_ctx_select", 0)
        return cid

    def _template(self, ncell):
        t = self._templates.get(ncell)
        if t is None:
            t = self._build(ncell)
            self._templates[ncell] = t
            self._tbytes = None
        return t

    def warm(self, ncell):
        """Build the template context of a grid size ahead of its first job."""
        self._template(ncell)

    def template_bytes(self):
        """Device bytes of the template contexts (cached until a template is
        built or destroyed: admission asks for it per job)."""
        if self._tbytes is None:
            self._tbytes = sum(self.held_bytes(t) for t in self._templates.values())
        return self._tbytes

    def open(self, spec: JobSpec):
        """A new job context: the mesh cloned from the size's template (device
        copies, no host work proportional to the mesh), the initial condition
        written on the device, the first dt reduced on the device."""
        mesh = self._mesh_for(spec.ncell)[0]
        tid = self._template(spec.ncell)
        cid = ctypes.c_int32()
        call("rh_ctx_create", ctypes.byref(cid))
        cid = cid.value
        call("rh_ctx_select", cid)
        try:
            call("rh_ctx_clone_mesh", tid)
            a = np.ascontiguousarray(spec.ic, dtype=np.float64)
            call("rh_init_ic", 1, mesh.noct, spec.ncell, spec.kind, ptr(a), self.p.gamma)
            g = np.zeros(3)
            call("rh_job_prepare", 1, mesh.noct, mesh.level, mesh.dx, self.p.gamma, self.p.smallr,
                 self.p.smallc ** 2, self.p.courant_factor, ptr(g), self.p.slope_type, self.p.riemann_id)
        except Exception:
            call("rh_ctx_destroy", cid)
            raise
        finally:
            call("rh_ctx_select", 0)
        return cid

    def close_all(self):
        """Destroy the template contexts."""
        for t in self._templates.values():
            call("rh_ctx_destroy", t)
        self._templates.clear()
        self._tbytes = None

    def step(self, handles, nsteps=1):
        ids = np.ascontiguousarray(handles, dtype=np.int32)
        call("rh_batch_step", ptr(ids, _lib._ip), len(ids), self.precision, nsteps)

    def sync(self):
        call("rh_device_sync")

    @staticmethod
    def _status(out):
        return {"t": out[0], "dt_next": out[1], "mass": out[2], "ekin": out[3], "eint": out[4]}

    def status(self, h):
        out = np.zeros(7)
        call("rh_job_status", h, ptr(out))
        return self._status(out)

    def fence(self, handles):
        """A fence after the work enqueued so far, holding the statuses of the
        jobs handles at that point (rh_fence_record); returns (fence, n)."""
        ids = np.ascontiguousarray(handles, dtype=np.int32)
        f = ctypes.c_int32()
        call("rh_fence_record", ptr(ids, _lib._ip), len(ids), ctypes.byref(f))
        return f.value, len(ids)

    def wait(self, fence):
        """Block until the fence has passed; the statuses it holds."""
        f, n = fence
        out = np.zeros((max(n, 1), 7))
        call("rh_fence_wait", f, ptr(out))
        return [self._status(out[j]) for j in range(n)]

    def state(self, h, spec: JobSpec):
        """Final conservative cells (5, N, N, N) of the job."""
        mesh = self._mesh_for(spec.ncell)[0]
        call("rh_ctx_select", h)
        try:
            u = np.zeros((mesh.noct, NVAR, 8))
            call("rh_transfer_grid_host", ptr(u), None, mesh.noct, NVAR, 8)
        finally:
            call("rh_ctx_select", 0)
        return mesh.to_cells(u)

    def initial_state(self, h, spec: JobSpec):
        return self.state(h, spec)

    def held_bytes(self, h):
        b = ctypes.c_int64()
        call("rh_ctx_footprint", h, ctypes.byref(b))
        return b.value

    def close(self, h):
        call("rh_ctx_destroy", h)


# ---------------------------------------------------------------------------
# Request loop
# ---------------------------------------------------------------------------
@dataclass
class JobResult:
    job_id: int
    ncell: int
    arrival: float
    admitted: float
    finished: float
    status: dict
    state: object = None

    @property
    def latency(self):
        return self.finished - self.arrival


class EnsembleServer:
    """Continuous batching of independent jobs over one engine.

    max_slots bounds the jobs resident at once; hbm_budget bounds the sum of
    their device footprints (bytes).  Admission is first come, first served:
    the head of the queue waits for room rather than being overtaken."""

    def __init__(self, engine, max_slots=64, hbm_budget=96 << 30, keep_states=False, clock=time.perf_counter,
                 sleep=time.sleep, depth=2):
        self.engine = engine
        self.depth = max(1, int(depth))   # ticks enqueued and not yet retired, at most
        self.max_slots = max_slots
        self.hbm_budget = hbm_budget
        self.keep_states = keep_states
        self.clock = clock
        self.sleep = sleep
        self.trace = []       # (t, n_active, bytes_in_use) after each tick

    def run(self, specs, realtime=True):
        """Serve the stream; returns {job_id: JobResult}.  realtime=False admits
        every job at t=0 regardless of its arrival time (closed batch).

        Ticks stay in flight: the loop enqueues tick k (the jobs' step counts
        are known on the host, so the jobs of tick k are too), then waits for
        tick k-depth+1 and retires the jobs that finished with it, then admits
        new jobs, all while the device runs the ticks still queued.  A job
        admitted after tick k was enqueued joins tick k+1."""
        eng = self.engine
        pending = deque(sorted(specs, key=lambda s: (s.arrival, s.job_id)))
        queue = deque()
        active = []           # [spec, handle, ticks_enqueued, admitted_at, bytes]
        resident = 0          # jobs holding device memory (active + finished, not yet retired)
        results = {}
        in_use = 0
        inflight = deque()    # (fence, jobs that finish with that tick), oldest first
        t0 = self.clock()
        ticks = 0
        while pending or queue or active or inflight:
            now = self.clock() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                queue.append(pending.popleft())
            # enqueue the next tick of every admitted job behind those in flight
            if active:
                eng.step([a[1] for a in active])
                ticks += 1
                done, still = [], []
                for a in active:
                    a[2] += 1
                    (done if a[2] >= a[0].nsteps else still).append(a)
                active = still
                inflight.append((eng.fence([a[1] for a in done]), done))
            # the oldest tick beyond the depth (every tick once nothing is left
            # to enqueue): wait for it and retire the jobs it finished
            if inflight and (len(inflight) >= self.depth or not active):
                f, done = inflight.popleft()
                sts = eng.wait(f)
                for a, st in zip(done, sts):
                    spec = a[0]
                    state = eng.state(a[1], spec) if self.keep_states else None
                    eng.close(a[1])
                    in_use -= a[4]
                    resident -= 1
                    results[spec.job_id] = JobResult(spec.job_id, spec.ncell, spec.arrival if realtime else 0.0,
                                                     a[3], self.clock() - t0, st, state)
                self.trace.append((self.clock() - t0, resident, in_use))
            # admissions overlap the ticks in flight
            while queue and resident < self.max_slots:
                need = eng.footprint(queue[0])
                reserved = eng.template_bytes()  # the per-size template meshes, resident beside the jobs
                if reserved + need > self.hbm_budget:
                    raise ValueError("job %d needs %d bytes, more than the HBM budget beside the %d bytes of "
                                     "templates" % (queue[0].job_id, need, reserved))
                if reserved + in_use + need > self.hbm_budget:
                    break
                spec = queue.popleft()
                h = eng.open(spec)
                active.append([spec, h, 0, self.clock() - t0, need])
                in_use += need
                resident += 1
            if not active and not inflight and pending:
                self.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
        self.ticks = ticks
        self.elapsed = self.clock() - t0
        return results


def latency_summary(results, specs):
    lat = np.array([r.latency for r in results.values()])
    cells = sum(s.cells * s.nsteps for s in specs)
    span = max(r.finished for r in results.values()) - min(s.arrival for s in specs)
    return {"jobs": len(results), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "cell_updates": int(cells),
            "span_s": float(span), "throughput_cell_updates_per_s": float(cells / span)}
