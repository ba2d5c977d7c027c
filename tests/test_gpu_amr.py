"""Coarse-fine path on the GPU against the oracle: ghost octs interpolated from
the parent level (interpol_hydro), zero fluxes at refined faces, refluxing into
the parent cells, restriction (upload_fine), all through the C-ABI.

Bar: 1e-12 per conserved variable on the level octs after each step (the
parent-cell reflux sums use FP64 atomics, so their order differs from the
serial reference in the last bits)."""
import numpy as np
import pytest

from ramses2_amd import HydroParams
from ramses2_amd.autotune import DEFAULT_SHAPE
from ramses2_amd.hydro import AmrHierarchy
from tests.amr_cases import (leaf_total, oracle_mesh, oracle_step, restrict_all, smooth_state, three_level_mesh,
                             two_level_mesh)
from tests.helpers import rel_err_per_var

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
