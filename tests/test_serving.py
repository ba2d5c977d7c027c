"""Ensemble serving on the CPU: the request stream and the scheduling of the
request loop, with an engine that counts steps instead of running kernels."""
import numpy as np

from ramses2_amd import ic
from ramses2_amd.serving import (KIND_BLAST, KIND_RIEMANN, EnsembleServer, ic_cells, latency_summary,
                                 make_stream, mean_job_cells)


def test_stream_is_seeded_and_in_range():
    a = make_stream(400, rate=50.0, seed=11)
    b = make_stream(400, rate=50.0, seed=11)
    assert [(s.arrival, s.ncell, s.kind, s.ic) for s in a] == [(s.arrival, s.ncell, s.kind, s.ic) for s in b]
    n = np.array([s.ncell for s in a])
    assert n.min() >= 32 and n.max() <= 128 and np.all(n % 2 == 0)
    # log-uniform sizes: the median near sqrt(32*128) = 64
    assert 50 <= np.median(n) <= 80
    kinds = np.array([s.kind for s in a])
    assert 0.4 < np.mean(kinds == KIND_BLAST) < 0.6
    arr = np.array([s.arrival for s in a])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    # Poisson stream: mean gap 1/rate
    assert abs(np.mean(np.diff(arr)) * 50.0 - 1.0) < 0.15
    for s in a:
        if s.kind == KIND_BLAST:
            assert 0.1 <= s.info["E"] <= 10 and 1e-6 <= s.info["p0"] <= 1e-2
        else:
            assert all(0.1 <= x <= 10 for x in s.info["rho"] + s.info["p"])
            assert s.info["axis"] in (0, 1, 2)
    assert mean_job_cells(a) == np.mean([s.ncell ** 3 for s in a])


def test_ic_cells_match_the_ic_module():
    """The serving initial conditions are ic.sedov_3d / ic.riemann_3d with the
    job's drawn parameters, bit for bit."""
    for spec in make_stream(12, rate=0.0, seed=3, nmin=8, nmax=24):
        c = ic_cells(spec)
        if spec.kind == KIND_BLAST:
            ref = ic.sedov_3d(spec.ncell, gamma=1.4, e_blast=spec.info["E"], p0=spec.info["p0"])
        else:
            n = spec.ncell
            rl, rr = spec.info["rho"]
            pl, pr = spec.info["p"]
            idx = np.arange(n)
            left = (idx + 0.5) / n < 0.5
            shape = [1, 1, 1]
            shape[2 - spec.info["axis"]] = n
            L = np.broadcast_to(left.reshape(shape), (n, n, n))
            from ramses2_amd.mesh import prim_to_cons
            z = np.zeros((n, n, n))
            ref = prim_to_cons(np.where(L, rl, rr), (z, z, z), np.where(L, pl, pr), 1.4)
        assert np.array_equal(c, ref)


class CountingEngine:
    """Stands in for HipEngine: records steps, footprints and the jobs resident."""

    def __init__(self, clock):
        self.clock = clock
        self.steps = {}
        self.open_set = set()
        self.bytes = {}
        self.max_open = 0
        self.max_bytes = 0
        self.order = []
        self.next = 1
        self.templates = 0
        self.next_fence = 0
        self.outstanding = {}
        self.max_outstanding = 0
        self.fenced = set()
        self.retirable = set()
        self.log = []

    def footprint(self, spec):
        return 1000 * spec.ncell ** 3

    def template_bytes(self):
        return self.templates

    def open(self, spec):
        h = self.next
        self.next += 1
        self.open_set.add(h)
        self.steps[h] = 0
        self.bytes[h] = self.footprint(spec)
        self.order.append(spec.job_id)
        self.log.append(("open", h))
        self.max_open = max(self.max_open, len(self.open_set))
        self.max_bytes = max(self.max_bytes, sum(self.bytes[x] for x in self.open_set))
        return h

    def step(self, handles, nsteps=1):
        assert len(set(handles)) == len(handles)
        for h in handles:
            assert h in self.open_set and h not in self.fenced
            self.steps[h] += nsteps
        self.clock.t += 0.001 * len(handles)
        self.log.append(("step", len(handles)))

    def fence(self, handles):
        """The statuses at this point of the stream; a job fenced here takes
        no further step."""
        f = self.next_fence
        self.next_fence += 1
        self.outstanding[f] = (list(handles), [{"steps": self.steps[h]} for h in handles])
        self.fenced.update(handles)
        self.max_outstanding = max(self.max_outstanding, len(self.outstanding))
        self.log.append(("fence", f))
        return f

    def wait(self, f):
        self.log.append(("wait", f))
        handles, sts = self.outstanding.pop(f)
        self.retirable.update(handles)
        return sts

    def sync(self):
        pass

    def status(self, h):
        return {"steps": self.steps[h]}

    def state(self, h, spec):
        return None

    def close(self, h):
        assert h in self.retirable      # its fence has passed
        self.open_set.remove(h)
        self.log.append(("close", h))


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-6)


def test_request_loop_schedules_fifo_within_slots_and_budget():
    specs = make_stream(60, rate=400.0, seed=5, nmin=8, nmax=32, nsteps=7)
    clock = FakeClock()
    eng = CountingEngine(clock)
    budget = 1000 * 3 * 32 ** 3
    srv = EnsembleServer(eng, max_slots=5, hbm_budget=budget, clock=clock, sleep=clock.sleep)
    res = srv.run(specs)
    assert sorted(res) == [s.job_id for s in specs]
    assert all(r.status["steps"] == 7 for r in res.values())
    assert eng.max_open <= 5 and eng.max_bytes <= budget
    assert eng.order == sorted(eng.order)          # first come, first served
    for s in specs:
        r = res[s.job_id]
        assert r.admitted >= s.arrival and r.finished > r.admitted
    summ = latency_summary(res, specs)
    assert summ["jobs"] == 60 and summ["p99_s"] >= summ["p50_s"] > 0
    assert not eng.open_set


def test_closed_batch_admits_everything_at_once():
    specs = make_stream(10, rate=1.0, seed=8, nmin=8, nmax=16, nsteps=3)
    clock = FakeClock()
    eng = CountingEngine(clock)
    srv = EnsembleServer(eng, max_slots=64, hbm_budget=1 << 40, clock=clock, sleep=clock.sleep)
    res = srv.run(specs, realtime=False)
    assert eng.max_open == 10 and srv.ticks == 3
    assert all(r.status["steps"] == 3 for r in res.values())
    assert not eng.outstanding and eng.max_outstanding == 2


def test_one_tick_stays_in_flight():
    """Tick k is enqueued before the loop waits for tick k-1, and a job is
    closed only after the fence of its last tick has passed."""
    specs = make_stream(2, rate=1.0, seed=8, nmin=8, nmax=16, nsteps=3)
    specs[1].nsteps = 2
    clock = FakeClock()
    eng = CountingEngine(clock)
    srv = EnsembleServer(eng, max_slots=64, hbm_budget=1 << 40, clock=clock, sleep=clock.sleep)
    res = srv.run(specs, realtime=False)
    assert eng.log == [("open", 1), ("open", 2),
                       ("step", 2), ("fence", 0),
                       ("step", 2), ("fence", 1), ("wait", 0),
                       ("step", 1), ("fence", 2), ("wait", 1), ("close", 2),
                       ("wait", 2), ("close", 1)]
    assert res[specs[0].job_id].status["steps"] == 3 and res[specs[1].job_id].status["steps"] == 2


def test_deeper_pipeline_keeps_more_ticks_queued():
    """depth=3: two ticks stay queued behind the one being retired; the jobs'
    results are the same."""
    specs = make_stream(2, rate=1.0, seed=8, nmin=8, nmax=16, nsteps=3)
    specs[1].nsteps = 2
    clock = FakeClock()
    eng = CountingEngine(clock)
    srv = EnsembleServer(eng, max_slots=64, hbm_budget=1 << 40, clock=clock, sleep=clock.sleep, depth=3)
    res = srv.run(specs, realtime=False)
    assert eng.log == [("open", 1), ("open", 2),
                       ("step", 2), ("fence", 0),
                       ("step", 2), ("fence", 1),
                       ("step", 1), ("fence", 2), ("wait", 0),
                       ("wait", 1), ("close", 2),
                       ("wait", 2), ("close", 1)]
    assert eng.max_outstanding == 3
    assert res[specs[0].job_id].status["steps"] == 3 and res[specs[1].job_id].status["steps"] == 2


def test_fifo_and_budget_hold_at_every_depth():
    for depth in (1, 2, 4):
        specs = make_stream(60, rate=400.0, seed=5, nmin=8, nmax=32, nsteps=7)
        clock = FakeClock()
        eng = CountingEngine(clock)
        budget = 1000 * 3 * 32 ** 3
        srv = EnsembleServer(eng, max_slots=5, hbm_budget=budget, clock=clock, sleep=clock.sleep, depth=depth)
        res = srv.run(specs)
        assert sorted(res) == [s.job_id for s in specs]
        assert all(r.status["steps"] == 7 for r in res.values())
        assert eng.max_open <= 5 and eng.max_bytes <= budget
        assert eng.order == sorted(eng.order)
        assert eng.max_outstanding <= depth and not eng.open_set


def test_admission_counts_the_resident_templates():
    """The per-size template meshes stay on the device beside the jobs, so
    admission keeps jobs + templates within the HBM budget."""
    specs = make_stream(40, rate=400.0, seed=6, nmin=8, nmax=32, nsteps=5)
    clock = FakeClock()
    eng = CountingEngine(clock)
    eng.templates = 1000 * 32 ** 3
    budget = 1000 * 3 * 32 ** 3
    srv = EnsembleServer(eng, max_slots=8, hbm_budget=budget, clock=clock, sleep=clock.sleep)
    res = srv.run(specs)
    assert sorted(res) == [s.job_id for s in specs]
    assert eng.max_bytes + eng.templates <= budget
