"""The benchmark configurations at their full sizes, checked through properties
that hold at any size (the oracle's CPU runs take minutes there, and the
parity tests cover the same kernels on smaller meshes against the oracle):
- periodic conservation of every conserved variable;
- mixed precision against FP64 (L1 per conserved variable, mass drift);
- the mirror and axis-permutation symmetry of a centred Sedov blast;
- composite (leaf-cell) conservation of a multi-level AMR Sedov step."""
import numpy as np
import pytest

from ramses2_amd import HydroParams, UniformLevel, UniformMesh, ic

pytestmark = pytest.mark.gpu


def _conserved(u0, u1, tol):
    for v in range(5):
        s0, s1 = u0[:, v].sum(), u1[:, v].sum()
        # momenta of a blast at rest sum to exactly zero at the start
        scale = max(np.abs(u0[:, v]).sum(), np.abs(u1[:, v]).sum())
        assert abs(s1 - s0) <= tol * scale, (v, s0, s1)


def test_turbulence_256():
    """256^3 Mach-5 turbulence, 10 device steps per precision (each with its own
    Courant dt): FP64 conserves to round-off; the mixed run stays within 1e-6
    (L1, relative) of FP64 and conserves mass as well."""
    print("case turbulence 256^3", flush=True)
    mesh = UniformMesh(3, 8)
    cells = ic.turbulence_3d(256)
    p = HydroParams(gamma=5.0 / 3.0, courant_factor=0.5, slope_type=2, riemann="hllc")
    a = UniformLevel(mesh, cells, p)
    u0 = a.state()
    a.run_device(10, precision=0)
    ua = a.state()
    _conserved(u0, ua, 1e-11)
    b = UniformLevel(mesh, cells, p)
    b.run_device(10, precision=1)
    ub = b.state()
    del cells
    for v in range(5):
        l1 = np.abs(ub[:, v] - ua[:, v]).sum() / np.abs(ua[:, v]).sum()
        assert l1 < 1e-6, (v, l1)
    assert abs(ub[:, 0].sum() - ua[:, 0].sum()) <= 1e-12 * ua[:, 0].sum()


def test_sedov_128_symmetry():
    """128^3 Sedov blast, 200 FP64 device steps: mass and energy conserved, and
    the blast stays symmetric under x -> -x and under the x <-> y, x <-> z
    exchanges (rounding differs between the axes only in the order of sums)."""
    print("case sedov 128^3", flush=True)
    mesh = UniformMesh(3, 7)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    lev = UniformLevel(mesh, ic.sedov_3d(128), p)
    u0 = lev.state()
    lev.run_device(200)
    u1 = lev.state()
    _conserved(u0, u1, 1e-11)
    c = mesh.to_cells(u1)  # [variable, z, y, x]
    scale = np.abs(c).max(axis=(1, 2, 3))
    m = c[:, :, :, ::-1].copy()
    m[1] = -m[1]
    assert np.all(np.abs(m - c).max(axis=(1, 2, 3)) <= 1e-9 * scale)
    for ax, var in ((2, 2), (1, 3)):  # x <-> y, x <-> z
        t = np.swapaxes(c, 3, ax).copy()
        t[[1, var]] = t[[var, 1]]
        assert np.all(np.abs(t - c).max(axis=(1, 2, 3)) <= 1e-8 * scale), ax
    # the blast has moved out of its initial cells
    rho = c[0]
    assert rho.max() > 2.0


def test_amr_sedov_levels_6_10():
    """The AMR Sedov case of the benchmark (levels 6-10, 614k octs), 5 batched
    FP64 steps: leaf-cell mass and energy conserved through reflux and
    restriction; the mixed run within 1e-6 of FP64 (L1 over all octs)."""
    print("case amr sedov 6-10", flush=True)
    from ramses2_amd.amr_build import sedov_amr
    from ramses2_amd.hydro import AmrHierarchy
    from tests.amr_cases import leaf_total

    am, u0 = sedov_amr(levelmin=6, levelmax=10, radius=0.3)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    m0, e0 = leaf_total(am, u0, 0), leaf_total(am, u0, 4)
    a = AmrHierarchy(am, u0.copy(), p)
    a.run_device(5, precision=0, use_graph=True)
    ua = a.state()
    assert abs(leaf_total(am, ua, 0) - m0) <= 1e-11 * m0
    assert abs(leaf_total(am, ua, 4) - e0) <= 1e-11 * e0
    b = AmrHierarchy(am, u0.copy(), p)
    b.run_device(5, precision=1, use_graph=True)
    ub = b.state()
    for v in (0, 4):
        l1 = np.abs(ub[:, v] - ua[:, v]).sum() / np.abs(ua[:, v]).sum()
        assert l1 < 1e-6, (v, l1)
