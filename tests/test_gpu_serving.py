"""Ensemble serving on the GPU: device initial conditions, a served job against
the CPU oracle on a box that is not a power of two, results alone vs in the
mix, and the admission footprint."""
import numpy as np
import pytest

from ramses2_amd import HydroParams, UniformMesh
from ramses2_amd.serving import EnsembleServer, HipEngine, JobSpec, KIND_BLAST, KIND_RIEMANN, ic_cells, make_stream
from tests.helpers import oracle_params, oracle_uniform, rel_err_per_var

pytestmark = pytest.mark.gpu

P = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")


def _spec(kind, n, nsteps=3):
    for s in make_stream(40, rate=0.0, seed=77, nmin=n, nmax=n, nsteps=nsteps):
        if s.kind == kind:
            return s
    raise AssertionError("no job of that kind")


@pytest.mark.parametrize("kind", [KIND_BLAST, KIND_RIEMANN])
def test_device_ic_matches_host(kind):
    spec = _spec(kind, 20)
    eng = HipEngine(P)
    h = eng.open(spec)
    try:
        got = eng.initial_state(h, spec)
    finally:
        eng.close(h)
    assert np.array_equal(got, ic_cells(spec))


@pytest.mark.parametrize("kind", [KIND_BLAST, KIND_RIEMANN])
def test_served_job_matches_oracle(kind):
    """3 Courant steps of a 20^3 job (10 octs per dimension, a periodic box
    inside the 16^3-oct level) against the oracle on the same box."""
    spec = _spec(kind, 20, nsteps=3)
    eng = HipEngine(P, precision=0)
    srv = EnsembleServer(eng, keep_states=True)
    res = srv.run([spec], realtime=False)
    got = res[spec.job_id].state
    mesh = UniformMesh.for_cells(3, 20)
    u = mesh.from_cells(ic_cells(spec))
    om = oracle_uniform(mesh, u)
    po = oracle_params(P)
    t = 0.0
    for _ in range(3):
        dt = om.courant_fine(mesh.level, po)[4]
        om.set_unew(mesh.level)
        om.godunov_fine(mesh.level, dt, po)
        om.set_uold(mesh.level)
        t += dt
    err = rel_err_per_var(mesh.from_cells(got), om.uold)
    assert np.all(err < 1e-12), err
    assert abs(res[spec.job_id].status["t"] - t) <= 1e-12 * t


@pytest.mark.parametrize("precision", [0, 1])
def test_result_same_alone_and_in_the_mix(precision):
    """Each job of a mixed batch (different sizes, both kinds, admitted while
    others run) ends bit for bit where it ends when served alone."""
    specs = make_stream(8, rate=0.0, seed=123, nmin=16, nmax=40, nsteps=6)
    # staggered: at most 3 resident and 4..7 steps each, so jobs join batches
    # that are already running
    for s in specs:
        s.nsteps = 4 + s.job_id % 4
    eng = HipEngine(P, precision=precision)
    mix = EnsembleServer(eng, max_slots=3, keep_states=True).run(specs, realtime=False)
    for s in specs:
        alone = EnsembleServer(HipEngine(P, precision=precision), keep_states=True).run(
            [JobSpec(s.job_id, 0.0, s.ncell, s.kind, s.ic, s.nsteps, s.info)], realtime=False)
        a, m = alone[s.job_id], mix[s.job_id]
        assert np.array_equal(a.state, m.state), s
        assert a.status["t"] == m.status["t"]


def test_footprint_estimate_bounds_allocation():
    eng = HipEngine(P)
    for n in (16, 34, 64):
        spec = _spec(KIND_BLAST, n)
        est = eng.footprint(spec)
        h = eng.open(spec)
        try:
            held = eng.held_bytes(h)
        finally:
            eng.close(h)
        assert 0.9 * est <= held <= est, (n, held, est)


def test_fence_holds_the_statuses_of_its_point_in_the_stream():
    """A fence recorded between two ticks holds the statuses after the first,
    bit for bit what rh_job_status reads after a synchronisation there; a
    fence of no jobs only orders."""
    specs = make_stream(3, rate=0.0, seed=7, nmin=16, nmax=24, nsteps=2)
    eng = HipEngine(P)
    hs = [eng.open(s) for s in specs]
    try:
        eng.step(hs)
        eng.sync()
        after1 = [eng.status(h) for h in hs]
        eng.step(hs)
        eng.sync()
        after2 = [eng.status(h) for h in hs]
    finally:
        for h in hs:
            eng.close(h)
    hs = [eng.open(s) for s in specs]
    try:
        eng.step(hs)
        f1 = eng.fence(hs[::-1])
        eng.step(hs)
        f2 = eng.fence([])
        f3 = eng.fence(hs[1:])
        got1 = eng.wait(f1)
        assert eng.wait(f2) == []
        got3 = eng.wait(f3)
    finally:
        for h in hs:
            eng.close(h)
    assert got1 == after1[::-1]
    assert got3 == after2[1:]
    assert all(a["t"] > b["t"] > 0 for a, b in zip(after2, after1))
