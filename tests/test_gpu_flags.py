"""Refinement flags on the GPU (rh_hydro_flag: hydro_flag with hydro_refine's
density and pressure criteria) against the oracle's hydro_flag, cell for cell,
on the same device state.  The oracle's flags are pinned to the reference's
own hydro_flag (tests/test_oracle_reflevel.py)."""
import numpy as np
import pytest

from ramses2_amd import HydroParams, UniformLevel, UniformMesh, ic
from ramses2_amd.hydro import AmrHierarchy
from tests.amr_cases import low_face_mesh, oracle_mesh, restrict_all, smooth_state, three_level_mesh, two_level_mesh
from tests.helpers import oracle_uniform

pytestmark = pytest.mark.gpu

P = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")


def check_amr(am, gpu, crit, periodic=(1, 1, 1), btype=((2, 2), (2, 2), (2, 2))):
    om = oracle_mesh(am, gpu.state(), 0, 1, periodic=periodic, bound_type=btype)
    total = 0
    for l in am.levels:
        fg, ng = gpu.hh.hydro_flag(am.head[l], am.noct[l], P, *crit)
        fo, no = om.hydro_flag(l, P.gamma, P.smallr, *crit)
        fo = fo[am.head[l] - 1:am.head[l] - 1 + am.noct[l]]
        assert np.array_equal(fg, fo), (l, np.argwhere(fg != fo)[:5])
        assert ng == no
        total += ng
    return total


@pytest.mark.parametrize("crit", [(0.1, 0.1), (0.05, -1.0), (-1.0, 0.3)])
@pytest.mark.parametrize("mesh", ["two", "three"])
def test_amr_flags_match_oracle(mesh, crit):
    am = two_level_mesh() if mesh == "two" else three_level_mesh()
    u = smooth_state(am, 5, blast=True)
    om = oracle_mesh(am, u, 0, 1)
    restrict_all(am, om)
    print("case", mesh, crit, flush=True)
    gpu = AmrHierarchy(am, om.uold.copy(), P)
    for _ in range(2):
        gpu.step()
    n = check_amr(am, gpu, crit)
    assert 0 < n < am.ngrid * 8


def test_low_face_flags_match_oracle():
    """Non-periodic in x, refined block at the low face: neighbours beyond the
    face come from the boundary octs of the level or of the coarser level."""
    am = low_face_mesh(0)
    u = smooth_state(am, 21, blast=True)
    bt = ((2, 2), (2, 2), (2, 2))
    om = oracle_mesh(am, u, 0, 1, periodic=(0, 1, 1), bound_type=bt)
    restrict_all(am, om)
    gpu = AmrHierarchy(am, om.uold.copy(), P, periodic=(0, 1, 1))
    gpu.step()
    n = check_amr(am, gpu, (0.1, 0.1), periodic=(0, 1, 1), btype=bt)
    assert n > 0


def test_uniform_sedov_flags_match_oracle():
    mesh = UniformMesh(3, 5)
    lev = UniformLevel(mesh, ic.sedov_3d(32), P)
    lev.run_device(20)
    om = oracle_uniform(mesh, lev.state())
    fg, ng = lev.hh.hydro_flag(1, mesh.noct, P, 0.1, 0.1)
    fo, no = om.hydro_flag(5, P.gamma, P.smallr, 0.1, 0.1)
    assert np.array_equal(fg, fo)
    assert ng == no > 0
    # flags are OR-ed: a second pass with a stricter criterion keeps them
    fg2, ng2 = lev.hh.hydro_flag(1, mesh.noct, P, 0.5, 0.5, reset=False)
    assert np.array_equal(fg2, fg) and ng2 == ng
