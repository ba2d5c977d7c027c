"""Coarse-fine path on the GPU against the oracle: ghost octs interpolated from
the parent level (interpol_hydro), zero fluxes at refined faces, refluxing into
the parent cells, restriction (upload_fine), all through the C-ABI.

Bar: 1e-12 per conserved variable on the level octs after each step (the
parent-cell reflux sums use FP64 atomics, so their order differs from the
serial reference in the last bits), and 1e-10 cell by cell (relative to each
cell's value, floored at 1e-6 of the variable's largest magnitude)."""
import numpy as np
import pytest

from ramses2_amd import HydroParams
from ramses2_amd.autotune import DEFAULT_SHAPE
from ramses2_amd.hydro import AmrHierarchy
from tests.amr_cases import (leaf_total, oracle_mesh, oracle_step, restrict_all, smooth_state, three_level_mesh,
                             two_level_mesh)
from tests.helpers import oracle_params, rel_err_per_cell, rel_err_per_var

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("interp", [(0, 1), (0, 2), (1, 2)])
@pytest.mark.parametrize("riemann", ["hllc", "llf"])
def test_two_level_steps(interp, riemann):
    am = two_level_mesh()
    u = smooth_state(am, 2, blast=True)
    om = oracle_mesh(am, u, *interp)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann=riemann)
    print("case", interp, riemann, flush=True)
    gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=interp[0], interpol_type=interp[1])
    for step in range(3):
        dt_o = oracle_step(am, om, p)
        dt_g = gpu.step()
        assert abs(dt_g - dt_o) <= 1e-14 * dt_o
        err = rel_err_per_var(gpu.state(), om.uold)
        assert np.all(err < 1e-12), (step, err)
        cell = rel_err_per_cell(gpu.state(), om.uold)
        print("per-cell relative error", step, cell)
        assert np.all(cell < 1e-10), (step, cell)


def test_two_level_conservation_gpu():
    am = two_level_mesh(offset=(5, 3, 1))
    u = smooth_state(am, 4)
    om = oracle_mesh(am, u, 0, 1)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=1, riemann="hllc")
    gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=0, interpol_type=1)
    u0 = gpu.state()
    for _ in range(4):
        gpu.step()
    u1 = gpu.state()
    for v in range(5):
        t0, t1 = leaf_total(am, u0, v), leaf_total(am, u1, v)
        assert abs(t1 - t0) <= 1e-13 * leaf_total(am, np.abs(u1), v), v


def test_two_level_mixed_precision():
    am = two_level_mesh()
    u = smooth_state(am, 2, blast=True)
    om = oracle_mesh(am, u, 0, 2)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=0, interpol_type=2, precision=1)
    dt = oracle_step(am, om, p)
    gpu.step(dt=dt)
    err = rel_err_per_var(gpu.state(), om.uold)
    assert np.all(err < 2e-6), err


@pytest.mark.parametrize("mesh", ["two", "three"])
@pytest.mark.parametrize("graph", [False, True])
def test_batched_levels_match_oracle(mesh, graph):
    """rh_amr_run: every level in one work-list launch per step, refluxes
    folded after the launch, restriction and the Courant pass on the device,
    eager or replayed as a hipGraph, against the oracle's amr_step."""
    am = two_level_mesh() if mesh == "two" else three_level_mesh()
    u = smooth_state(am, 6, blast=True)
    om = oracle_mesh(am, u, 0, 2)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    print("case", mesh, graph, flush=True)
    gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=0, interpol_type=2)
    t = 0.0
    for _ in range(4):
        t += oracle_step(am, om, p)
    dt, tt, sums, ms = gpu.run_device(4, use_graph=graph)
    err = rel_err_per_var(gpu.state(), om.uold)
    assert np.all(err < 1e-12), err
    assert abs(tt - t) <= 1e-13 * t
    mass = leaf_total(am, om.uold, 0)
    assert abs(sums[0] - mass) <= 1e-12 * abs(mass)
    # a fifth step from where the run stopped (odd count: the arrays swap back)
    dt_o = oracle_step(am, om, p)
    assert abs(dt - dt_o) <= 1e-14 * dt_o
    gpu.run_device(1)
    assert np.all(rel_err_per_var(gpu.state(), om.uold) < 1e-12)


def test_batched_levels_mixed_precision():
    am = three_level_mesh()
    u = smooth_state(am, 9, blast=True)
    om = oracle_mesh(am, u, 0, 2)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=0, interpol_type=2)
    oracle_step(am, om, p)
    gpu.run_device(1, precision=1)
    err = rel_err_per_var(gpu.state(), om.uold)
    assert np.all(err < 2e-6), err


def test_batched_mixed_courant_in_fp32():
    """The mixed batched step reduces each leaf cell's dt from an FP32 signal
    speed: its next dt agrees with the oracle's FP64 courant_fine of the state
    it leaves to 1e-5, and its leaf mass sum is the FP64 one."""
    am = three_level_mesh()
    u = smooth_state(am, 9, blast=True)
    om = oracle_mesh(am, u, 0, 2)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=0, interpol_type=2)
    dt_next, t, sums, _ = gpu.run_device(2, precision=1)
    u2 = gpu.state()
    o2 = oracle_mesh(am, u2, 0, 2)
    po = oracle_params(p)
    dt_o = min(min(am.dx(l) / p.smallc, o2.courant_fine(l, po)[4]) for l in am.levels)
    assert abs(dt_next - dt_o) <= 1e-5 * dt_o, (dt_next, dt_o)
    mass = leaf_total(am, u2, 0)
    assert abs(sums[0] - mass) <= 1e-12 * abs(mass), (sums[0], mass)


def test_batched_levels_with_per_level_brick_shapes():
    """Levels built with different brick shapes: one work-list launch per
    shape inside the batched step, same answer as the oracle."""
    am = three_level_mesh()
    u = smooth_state(am, 12, blast=True)
    om = oracle_mesh(am, u, 0, 2)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    lo = am.levelmin
    try:
        gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=0, interpol_type=2,
                           bricks={lo: (7, 7, 8), lo + 1: (3, 3, 4), lo + 2: (15, 7, 8)})
        for _ in range(3):
            oracle_step(am, om, p)
        gpu.run_device(3, use_graph=True)
    finally:
        from ramses2_amd.hydro import HipHydro
        HipHydro.set_brick_shape(*DEFAULT_SHAPE)
    err = rel_err_per_var(gpu.state(), om.uold)
    assert np.all(err < 1e-12), err


@pytest.mark.parametrize("btype", [2, 1])
def test_nonperiodic_low_face(btype):
    """A refined block touching the low face of a box non-periodic in x
    (zero-gradient or reflexive): ghost octs beside the face take parent
    cells from the coarse level's boundary octs (floor division of keys), and
    stencil octs beyond the face whose reference oct is missing are
    interpolated from those boundary octs."""
    from tests.amr_cases import low_face_mesh
    am = low_face_mesh(0)
    u = smooth_state(am, 21, blast=True)
    bt = ((btype, btype), (2, 2), (2, 2))
    om = oracle_mesh(am, u, 0, 1, periodic=(0, 1, 1), bound_type=bt)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    print("case", btype, flush=True)
    gpu = AmrHierarchy(am, om.uold.copy(), p, periodic=(0, 1, 1), bound_type=(btype, btype, 2, 2, 2, 2))
    for step in range(3):
        dt_o = oracle_step(am, om, p)
        dt_g = gpu.step()
        assert abs(dt_g - dt_o) <= 1e-14 * dt_o
        err = rel_err_per_var(gpu.state(), om.uold)
        assert np.all(err < 1e-12), (step, err)


def test_two_nonperiodic_dimensions_rejected():
    """3D boxes with two or more non-periodic dimensions have stencil octs
    beyond an edge of the box, which the reference builds from whatever its
    cache holds: the library refuses them."""
    from ramses2_amd._lib import HipError
    am = two_level_mesh()
    u = smooth_state(am, 1)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    with pytest.raises(HipError, match="non-periodic"):
        AmrHierarchy(am, u, p, periodic=(0, 0, 1))


@pytest.mark.parametrize("path", ["dropin", "batched"])
def test_deterministic_reflux_is_repeatable(path):
    """rh_set_deterministic(1): every coarse-fine face writes its own slot and
    the slots are added to the parent cell in a fixed order, so two runs from
    the same state give the same bits (with FP64 atomics they may not); the
    result still meets the oracle bar."""
    am = three_level_mesh()
    u = smooth_state(am, 13, blast=True)
    om = oracle_mesh(am, u, 0, 2)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    print("case", path, flush=True)
    states = []
    try:
        for _ in range(2):
            gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=0, interpol_type=2, deterministic=True)
            if path == "dropin":
                for _ in range(3):
                    gpu.step()
            else:
                gpu.run_device(3, use_graph=True)
            states.append(gpu.state())
    finally:
        from ramses2_amd._lib import call
        call("rh_set_deterministic", 0)
    assert np.array_equal(states[0], states[1])
    for _ in range(3):
        oracle_step(am, om, p)
    assert np.all(rel_err_per_var(states[0], om.uold) < 1e-12)


@pytest.mark.parametrize("deterministic", [False, True])
def test_two_level_diffusion(deterministic):
    """difmag = 0.1 on a two-level mesh through the per-level calls: the
    diffusion pass's fluxes are refluxed into the parent cells with the
    Godunov fluxes (atomically, or into the per-face slots)."""
    from ramses2_amd._lib import call
    am = two_level_mesh()
    u = smooth_state(am, 17, blast=True)
    om = oracle_mesh(am, u, 0, 1)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc", difmag=0.1)
    print("case", deterministic, flush=True)
    try:
        gpu = AmrHierarchy(am, om.uold.copy(), p, deterministic=deterministic)
        for step in range(2):
            dt_o = oracle_step(am, om, p)
            dt_g = gpu.step()
            assert abs(dt_g - dt_o) <= 1e-14 * dt_o
            assert np.all(rel_err_per_var(gpu.state(), om.uold) < 1e-12)
        u0 = gpu.state()
    finally:
        call("rh_set_difmag", 0.0)
        call("rh_set_deterministic", 0)
    mass = [leaf_total(am, x, 0) for x in (om.uold, u0)]
    assert abs(mass[1] - mass[0]) <= 1e-13 * abs(mass[0])


@pytest.mark.parametrize("shape", [(4, 4, 4), (3, 3, 4)])
@pytest.mark.parametrize("path", ["dropin", "batched", "batched_mixed"])
def test_lds_reflux_bricks_match_oracle(path, shape):
    """Bricks whose window fits the LDS reflux sums (4x4x4: 216 window octs,
    3x3x4: 150) add a brick's refluxes per ghost oct in LDS and flush them
    with one atomic per ghost oct and variable: the same answer as the
    oracle, through the per-level drop-in calls (sums into the parent cells)
    and the batched step (sums into the accumulator), three levels."""
    am = three_level_mesh()
    u = smooth_state(am, 5, blast=True)
    om = oracle_mesh(am, u, 0, 2)
    restrict_all(am, om)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    print("case", path, shape, flush=True)
    try:
        gpu = AmrHierarchy(am, om.uold.copy(), p, interpol_var=0, interpol_type=2,
                           bricks={l: shape for l in am.levels})
        if path == "dropin":
            for step in range(3):
                dt_o = oracle_step(am, om, p)
                dt_g = gpu.step()
                assert abs(dt_g - dt_o) <= 1e-14 * dt_o
                assert np.all(rel_err_per_var(gpu.state(), om.uold) < 1e-12), step
                assert np.all(rel_err_per_cell(gpu.state(), om.uold) < 1e-10), step
        elif path == "batched":
            for _ in range(3):
                oracle_step(am, om, p)
            gpu.run_device(3, use_graph=True)
            assert np.all(rel_err_per_var(gpu.state(), om.uold) < 1e-12)
            assert np.all(rel_err_per_cell(gpu.state(), om.uold) < 1e-10)
        else:
            oracle_step(am, om, p)
            gpu.run_device(1, precision=1)
            assert np.all(rel_err_per_var(gpu.state(), om.uold) < 2e-6)
    finally:
        from ramses2_amd.hydro import HipHydro
        HipHydro.set_brick_shape(*DEFAULT_SHAPE)
