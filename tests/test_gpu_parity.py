"""HIP path against the CPU oracle, through the C-ABI (include/ramses2_hip.h).

Bar: the FP64 kernel within 1e-12 of the oracle per conserved variable, measured
as max_cells|gpu-oracle| / max_cells|oracle| for each variable, on one step
with the same dt (FMA contraction is the only allowed difference), and within
TOLCELL of it cell by cell: |gpu-oracle| / max(|oracle|, 1e-6 max_cells|oracle|)
for every cell and variable.  The mixed kernel (FP32 reconstruction/fluxes,
FP64 accumulation) within 2e-6.
"""
import numpy as np
import pytest

from ramses2_amd import HydroParams, UniformLevel, UniformMesh, ic
from ramses2_amd.autotune import DEFAULT_SHAPE
from tests.helpers import oracle_params, oracle_uniform, rel_err_per_cell, rel_err_per_var

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOLCELL = 1e-11
TOLMIX = 2e-6


def _one_step_3d(level, cells, p, dt=None, precision=0, brick=None):
    mesh = UniformMesh(3, level)
    lev = None
    from ramses2_amd.hydro import HipHydro
    if brick is not None:
        HipHydro.set_brick_shape(*brick)
    lev = UniformLevel(mesh, cells, p)
    u0 = lev.state()
    om = oracle_uniform(mesh, u0)
    po = oracle_params(p)
    c_ora = om.courant_fine(level, po)
    c_gpu = lev.courant()
    if dt is None:
        dt = c_ora[4]
    lev.step(dt=dt, precision=precision)
    om.set_unew(level)
    om.godunov_fine(level, dt, po)
    return lev.state(), om.unew.copy(), c_gpu, c_ora, u0


@pytest.mark.parametrize("slope", [1, 2, 3])
@pytest.mark.parametrize("riemann", ["llf", "hll", "hllc"])
def test_godunov3d_random_fp64(slope, riemann):
    rng = np.random.default_rng(100 + slope * 7 + len(riemann))
    mesh_level = 4  # 16^3 cells, 8^3 octs: partial bricks and periodic wrap inside bricks
    cells = ic.random_state((16, 16, 16), rng, vscale=1.0)
    p = HydroParams(gamma=1.4, courant_factor=0.8, slope_type=slope, riemann=riemann)
    g, o, _, _, _ = _one_step_3d(mesh_level, cells, p)
    err = rel_err_per_var(g, o)
    assert np.all(err < TOL64), err
    cell = rel_err_per_cell(g, o)
    print("per-cell relative error", cell)
    assert np.all(cell < TOLCELL), cell


def test_godunov3d_floors_fp64():
    """Near-vacuum cells exercise the density/pressure floors of ctoprim, trace3d and HLLC."""
    rng = np.random.default_rng(7)
    cells = ic.random_state((16, 16, 16), rng, vscale=2.0, near_vacuum=0.05)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    g, o, _, _, _ = _one_step_3d(4, cells, p)
    err = rel_err_per_var(g, o)
    assert np.all(err < 1e-10), err


BRICKS = [(7, 7, 8), (7, 7, 16), (8, 6, 8), (8, 6, 16), (4, 4, 4), (4, 4, 8), (15, 7, 8), (15, 7, 16), (3, 3, 4),
          (1, 1, 1), (3, 3, 8), (7, 7, 4)]


@pytest.mark.parametrize("variant", [(2, "hllc"), (3, "llf"), (1, "hll")])
@pytest.mark.parametrize("brick", BRICKS)
def test_brick_shapes_agree(brick, variant):
    """Every compiled brick shape, in each kernel variant (HLLC; the 27-point
    slope's three-plane ring; the generic solver path), gives the oracle's
    answer."""
    rng = np.random.default_rng(11)
    cells = ic.random_state((16, 16, 16), rng)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=variant[0], riemann=variant[1])
    try:
        g, o, _, _, _ = _one_step_3d(4, cells, p, brick=brick)
    finally:
        from ramses2_amd.hydro import HipHydro
        HipHydro.set_brick_shape(*DEFAULT_SHAPE)
    assert np.all(rel_err_per_var(g, o) < TOL64)
    assert np.all(rel_err_per_cell(g, o) < TOLCELL), rel_err_per_cell(g, o)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("brick", BRICKS)
def test_fused_courant_every_shape(brick, precision):
    """The Courant reduction in the Godunov kernel's epilogue, for every brick
    shape (the 8x6, 4x4 and 1x1x1 workgroups end in a partial wave), steps with
    the dts of the separate Courant pass: the same time and state, bit for bit
    in FP64 (dt is a min, exact in any order), to the mixed tolerance in FP32
    (the epilogue's signal speed is FP32 there).  The returned dt differs by
    design: the separate pass reports the last step's dt, the epilogue the
    next one's; the conserved sums (mass, energy) agree."""
    from ramses2_amd.hydro import HipHydro
    mesh = UniformMesh(3, 4)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    cells = ic.sedov_3d(16)
    HipHydro.set_brick_shape(*brick)
    try:
        a = UniformLevel(mesh, cells, p)
        dta, ta, sa, _ = a.run_device(3, precision=precision, fused_courant=0)
        b = UniformLevel(mesh, cells, p)
        dtb, tb, sb, _ = b.run_device(3, precision=precision, fused_courant=1)
    finally:
        HipHydro.set_brick_shape(*DEFAULT_SHAPE)
    assert dta > 0 and dtb > 0 and np.isfinite(dtb) and abs(dtb - dta) < 0.05 * dta
    if precision == 0:
        assert tb == ta
        assert np.array_equal(b.state(), a.state())
    else:
        assert abs(tb - ta) <= 1e-5 * ta
        assert np.all(rel_err_per_var(b.state(), a.state()) < 1e-5)
    assert np.allclose(sb[:2], sa[:2], rtol=1e-12, atol=0)


def test_courant_fp64():
    rng = np.random.default_rng(3)
    cells = ic.random_state((16, 16, 16), rng, vscale=3.0)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    _, _, cg, co, _ = _one_step_3d(4, cells, p)
    assert abs(cg["dt"] - co[4]) <= 1e-14 * co[4]
    for k, i in (("mass", 0), ("ekin", 1), ("eint", 2)):
        assert abs(cg[k] - co[i]) <= 1e-13 * abs(co[i]), (k, cg[k], co[i])


def test_sedov_steps_fp64():
    """Sedov 32^3, 10 steps, each step compared with the oracle on the same input."""
    level = 5
    mesh = UniformMesh(3, level)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    lev = UniformLevel(mesh, ic.sedov_3d(32), p)
    po = oracle_params(p)
    for step in range(10):
        u0 = lev.state()
        om = oracle_uniform(mesh, u0)
        c = om.courant_fine(level, po)
        dt = c[4]
        lev.step(dt=dt)
        om.godunov_fine(level, dt, po)
        err = rel_err_per_var(lev.state(), om.unew)
        assert np.all(err < TOL64), (step, err)


def test_sod_1d_run():
    """Sod tube, 1024 cells, MinMod + HLLC, Courant 0.8, zero-gradient ends, to t=0.2:
    every step is compared with the oracle on the same input and dt, and the final
    states of the two independent runs (each with its own dt sequence) agree."""
    mesh = UniformMesh(1, 10)
    p = HydroParams(gamma=1.4, courant_factor=0.8, slope_type=1, riemann="hllc")
    cells = ic.sod_1d(1024)
    lev = UniformLevel(mesh, cells, p, periodic=False, bound_type=2)
    po = oracle_params(p)
    from ramses2_amd.hydro import newdt
    # independent oracle run
    om_run = oracle_uniform(mesh, mesh.from_cells(cells), periodic=False)
    t, dtold, nstep = 0.0, 0.0, 0
    while t < 0.2 and nstep < 2000:
        c = om_run.courant_fine(10, po)
        dt = newdt(mesh.dx, p, c[4], t, 0.2, 3, dtold)
        om_run.set_unew(10)
        om_run.godunov_fine(10, dt, po)
        om_run.set_uold(10)
        t += dt
        dtold = dt
        nstep += 1
    # GPU run, per-step check
    while lev.t < 0.2 and lev.nstep < 2000:
        u0 = lev.state()
        om = oracle_uniform(mesh, u0, periodic=False)
        c = om.courant_fine(10, po)
        cg = lev.courant()
        assert abs(cg["dt"] - c[4]) <= 1e-14 * c[4]
        dt = newdt(mesh.dx, p, cg["dt"], lev.t, 0.2, 3, lev.dtold)
        lev.step(dt=dt)
        om.godunov_fine(10, dt, po)
        err = rel_err_per_var(lev.state(), om.unew)
        assert np.all(err < TOL64), (lev.nstep, err)
    assert lev.nstep == nstep
    assert abs(lev.t - t) < 1e-12
    err = rel_err_per_var(lev.state(), om_run.uold)
    assert np.all(err < 1e-9), err


def test_mixed_precision_one_step():
    rng = np.random.default_rng(5)
    cells = ic.random_state((16, 16, 16), rng, vscale=1.0)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    g, o, _, _, u0 = _one_step_3d(4, cells, p, precision=1)
    # error of the update relative to the size of the update
    err = np.max(np.abs(g - o)) / np.max(np.abs(o - u0))
    assert err < 1e-5, err
    assert np.all(rel_err_per_var(g, o) < TOLMIX)


@pytest.mark.parametrize("fused", [0, 1])
def test_device_run_matches_dropin_calls(fused):
    """rh_level_run (dt on the device, set_unew/set_uold folded, hipGraph, the
    Courant pass separate or in the Godunov epilogue) gives the same state as
    the drop-in call sequence, bit for bit (dt is a min, exact in any order)."""
    level = 4
    mesh = UniformMesh(3, level)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    cells = ic.sedov_3d(16)
    a = UniformLevel(mesh, cells, p)
    for _ in range(6):
        a.step()
    ua = a.state()
    b = UniformLevel(mesh, cells, p)
    b.run_device(6, use_graph=False, time_kernel=True, fused_courant=fused)
    assert np.array_equal(b.state(), ua)
    c = UniformLevel(mesh, cells, p)
    c.run_device(6, use_graph=True, fused_courant=fused)
    assert np.array_equal(c.state(), ua)
    assert abs(c.t - a.t) <= 1e-15 * a.t


def test_conservation_periodic():
    """On a periodic box the FP64 step conserves mass and energy to round-off."""
    level = 5
    mesh = UniformMesh(3, level)
    p = HydroParams(gamma=5.0 / 3.0, courant_factor=0.5, slope_type=2, riemann="hllc")
    lev = UniformLevel(mesh, ic.turbulence_3d(32, seed=3), p)
    u0 = lev.state()
    lev.run_device(5)
    u1 = lev.state()
    for v in (0, 1, 2, 3, 4):
        s0, s1 = u0[:, v].sum(), u1[:, v].sum()
        scale = np.abs(u0[:, v]).sum()
        assert abs(s1 - s0) <= 1e-12 * scale, (v, s0, s1)


@pytest.mark.parametrize("slope", [1, 2])
def test_constant_gravity_fp64(slope):
    """A constant gravity vector: ctoprim's g dt/2 velocity kick and cmpdt's
    gravity term (hydro/umuscl.f90:1739-1901, hydro/godunov_utils.f90:228-337),
    through the drop-in calls, against the oracle (itself pinned to the
    reference's godunov_fine with gravity)."""
    from oracle import oracle as O
    rng = np.random.default_rng(21 + slope)
    mesh = UniformMesh(3, 4)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=slope, riemann="hllc")
    g = (0.7, -0.4, 0.25)
    lev = UniformLevel(mesh, ic.random_state((16, 16, 16), rng), p)
    u0 = lev.state()
    om = O.OracleMesh(3, mesh.ckey, np.full(mesh.noct, 4), np.zeros((mesh.noct, 8)), u0, [0, 0, 0, 1],
                      [0, 0, 0, mesh.noct], levelmin=4, gravity=g)
    po = oracle_params(p)
    c_o = om.courant_fine(4, po)
    c_g = lev.hh.cmpdt(1, mesh.noct, mesh.dx, p, gravity=g)
    assert abs(c_g["dt"] - c_o[4]) <= 1e-14 * c_o[4]
    dt = c_o[4]
    lev.hh.set_unew(1, mesh.noct)
    lev.hh.godunov(1, mesh.noct, 4, 4, 4, dt, mesh.dx, p, gravity=g)
    lev.hh.set_uold(1, mesh.noct)
    om.set_unew(4)
    om.godunov_fine(4, dt, po)
    err = rel_err_per_var(lev.state(), om.unew)
    assert np.all(err < TOL64), err
    assert np.all(rel_err_per_cell(lev.state(), om.unew) < TOLCELL)


@pytest.mark.parametrize("riemann", ["hllc", "llf"])
def test_artificial_diffusion_fp64(riemann):
    """difmag > 0: the diffusive fluxes of cmpdivu/cmpdiff (hydro/umuscl.f90:
    2392-2540) added after the Godunov launch, with constant gravity (whose
    half kick enters the node divergences), through the drop-in calls and
    through rh_level_run, against the oracle (pinned to the reference's
    godunov_fine with difmag = 0.1)."""
    from oracle import oracle as O
    rng = np.random.default_rng(33)
    mesh = UniformMesh(3, 4)
    p = HydroParams(gamma=5.0 / 3.0, courant_factor=0.5, slope_type=2, riemann=riemann, difmag=0.1)
    g = (0.3, 0.0, -0.2)
    cells = ic.random_state((16, 16, 16), rng, vscale=2.0)
    po = oracle_params(p)
    try:
        lev = UniformLevel(mesh, cells, p)
        u0 = lev.state()
        om = O.OracleMesh(3, mesh.ckey, np.full(mesh.noct, 4), np.zeros((mesh.noct, 8)), u0, [0, 0, 0, 1],
                          [0, 0, 0, mesh.noct], levelmin=4, gravity=g)
        dt = om.courant_fine(4, po)[4]
        lev.hh.set_unew(1, mesh.noct)
        lev.hh.godunov(1, mesh.noct, 4, 4, 4, dt, mesh.dx, p, gravity=g)
        lev.hh.set_uold(1, mesh.noct)
        om.set_unew(4)
        om.godunov_fine(4, dt, po)
        assert np.all(rel_err_per_var(lev.state(), om.unew) < TOL64)
        assert np.all(rel_err_per_cell(lev.state(), om.unew) < TOLCELL)
        # without diffusion the result differs: the pass does something
        q = HydroParams(gamma=p.gamma, courant_factor=0.5, slope_type=2, riemann=riemann)
        lev0 = UniformLevel(mesh, cells, q)
        lev0.hh.set_unew(1, mesh.noct)
        lev0.hh.godunov(1, mesh.noct, 4, 4, 4, dt, mesh.dx, q, gravity=g)
        assert np.max(np.abs(lev0.state() - om.unew)) > 1e-6 * np.max(np.abs(om.unew))
        # the device-resident step (Courant pass separate, dt on the device)
        lev2 = UniformLevel(mesh, cells, p)
        om2 = O.OracleMesh(3, mesh.ckey, np.full(mesh.noct, 4), np.zeros((mesh.noct, 8)), u0, [0, 0, 0, 1],
                           [0, 0, 0, mesh.noct], levelmin=4, gravity=g)
        for _ in range(2):
            d = om2.courant_fine(4, po)[4]
            om2.set_unew(4)
            om2.godunov_fine(4, d, po)
            om2.set_uold(4)
        lev2.run_device(2, gravity=g)
        assert np.all(rel_err_per_var(lev2.state(), om2.uold) < TOL64)
    finally:
        from ramses2_amd._lib import call
        call("rh_set_difmag", 0.0)


@pytest.mark.parametrize("precision", [0, 1])
def test_specialized_body_gives_the_same_bits(precision):
    """The brick body specialised for periodic levels without refined cells or
    refluxes (the default for such levels) against the general body
    (rh_set_specialized(0)) on the same state: bit for bit, three steps of the
    device-resident level stepping, FP64 and mixed."""
    from ramses2_amd._lib import call
    mesh = UniformMesh(3, 5)
    cells = ic.turbulence_3d(32, seed=17)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    out = []
    for spec in (1, 0):
        call("rh_set_specialized", spec)
        try:
            lev = UniformLevel(mesh, cells, p)
            dt, t, sums, _ = lev.run_device(3, precision=precision)
            out.append((lev.state(), dt, t))
        finally:
            call("rh_set_specialized", 1)
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2]
