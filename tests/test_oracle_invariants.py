"""Closed-form and invariant checks of the CPU oracle (pins what the golden
vectors leave open: the level driver godfine1, courant_fine, boundaries).

Each check is chosen so that a plausible slip fails it: a sign or index error
in the gather breaks the mirror/permutation symmetries, a dropped flux term
breaks conservation or the uniform-state test, a wrong wave speed breaks the
exact-flux and upwind checks, and the Sod run is held against Toro's exact
Riemann solution.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ramses2_amd import HydroParams, UniformMesh, ic
from tests.helpers import oracle_params, oracle_uniform


def euler_flux(q, gamma):
    r, u, v, w, p = q
    E = p / (gamma - 1) + 0.5 * r * (u * u + v * v + w * w)
    return np.array([r * u, r * u * u + p, r * u * v, r * u * w, (E + p) * u])


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_identical_states_give_exact_flux(solver):
    rng = np.random.default_rng(solver)
    p = O.make_params(gamma=1.4)
    for _ in range(50):
        q = np.array([np.exp(rng.normal()), *rng.normal(0, 2, 3), np.exp(rng.normal())])
        f = O.riemann(solver, q, q, p)[:5]
        ref = euler_flux(q, 1.4)
        assert np.allclose(f, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("solver", [2, 3])
def test_supersonic_upwind(solver):
    """Both states moving right faster than sound: HLL/HLLC return the left flux."""
    p = O.make_params(gamma=1.4)
    ql = np.array([1.0, 5.0, 0.3, -0.2, 1.0])
    qr = np.array([0.5, 6.0, 0.1, 0.4, 0.7])
    f = O.riemann(solver, ql, qr, p)[:5]
    assert np.allclose(f, euler_flux(ql, 1.4), rtol=1e-14)
    # and mirrored: moving left, right flux
    ql2 = ql * np.array([1, -1, 1, 1, 1])
    qr2 = qr * np.array([1, -1, 1, 1, 1])
    f2 = O.riemann(solver, qr2, ql2, p)[:5]
    assert np.allclose(f2, euler_flux(ql2, 1.4), rtol=1e-14)


def test_hllc_resolves_contact():
    """A stationary contact (u=0, equal pressure) carries no mass/energy flux under HLLC."""
    p = O.make_params(gamma=1.4)
    f = O.riemann(3, np.array([1.0, 0, 0.2, 0.1, 2.0]), np.array([0.1, 0, -0.3, 0.5, 2.0]), p)
    assert f[0] == 0 and f[4] == 0 and f[2] == 0 and f[3] == 0
    assert f[1] == 2.0


def test_uniform_state_is_steady():
    mesh = UniformMesh(3, 4)
    cells = np.zeros((5, 16, 16, 16))
    cells[0] = 1.3
    cells[1] = 1.3 * 0.7
    cells[2] = -1.3 * 0.2
    cells[3] = 1.3 * 0.1
    cells[4] = 2.5 / 0.4 + 0.5 * 1.3 * (0.49 + 0.04 + 0.01)
    u = mesh.from_cells(cells)
    om = oracle_uniform(mesh, u)
    po = oracle_params(HydroParams(slope_type=2, riemann="hllc"))
    om.godunov_fine(4, 1e-3, po)
    assert np.array_equal(om.unew, u)


def _step(mesh, cells, p, dt):
    om = oracle_uniform(mesh, mesh.from_cells(cells))
    om.godunov_fine(mesh.level, dt, oracle_params(p))
    return mesh.to_cells(om.unew)


@pytest.mark.parametrize("riemann", ["llf", "hllc"])
def test_mirror_symmetry(riemann):
    """Reflecting x -> -x (and rho u -> -rho u) commutes with the step."""
    rng = np.random.default_rng(2)
    mesh = UniformMesh(3, 4)
    cells = ic.random_state((16, 16, 16), rng)
    p = HydroParams(gamma=1.4, slope_type=2, riemann=riemann)
    dt = 2e-3
    a = _step(mesh, cells, p, dt)
    refl = cells[:, :, :, ::-1].copy()
    refl[1] *= -1
    b = _step(mesh, refl, p, dt)
    b = b[:, :, :, ::-1].copy()
    b[1] *= -1
    for v in range(5):
        assert np.max(np.abs(a[v] - b[v])) <= 1e-13 * np.max(np.abs(a[v])), v


def test_axis_permutation_symmetry():
    """Exchanging the x and z axes (and the x/z momenta) commutes with the step."""
    rng = np.random.default_rng(4)
    mesh = UniformMesh(3, 4)
    cells = ic.random_state((16, 16, 16), rng)
    p = HydroParams(gamma=1.4, slope_type=2, riemann="hllc")
    dt = 2e-3
    a = _step(mesh, cells, p, dt)
    perm = np.transpose(cells, (0, 3, 2, 1)).copy()
    perm[[1, 3]] = perm[[3, 1]]
    b = _step(mesh, perm, p, dt)
    b = np.transpose(b, (0, 3, 2, 1)).copy()
    b[[1, 3]] = b[[3, 1]]
    for v in range(5):
        assert np.max(np.abs(a[v] - b[v])) <= 1e-13 * np.max(np.abs(a[v])), v


def test_periodic_conservation():
    rng = np.random.default_rng(6)
    mesh = UniformMesh(3, 4)
    cells = ic.random_state((16, 16, 16), rng, vscale=2.0)
    p = HydroParams(gamma=1.4, slope_type=1, riemann="hllc")
    b = _step(mesh, cells, p, 1e-3)
    for v in range(5):
        assert abs(b[v].sum() - cells[v].sum()) <= 1e-13 * np.abs(cells[v]).sum()


def test_courant_closed_form():
    """cmpdt of a uniform moving gas: dt = 2 C dx / (ctot (1 + sqrt(1 + 2 C 1e-4)))."""
    mesh = UniformMesh(3, 3)
    rho, vel, pr, gamma, C = 1.2, (0.3, -0.4, 0.5), 0.9, 1.4, 0.5
    cells = np.zeros((5, 8, 8, 8))
    cells[0] = rho
    for d in range(3):
        cells[1 + d] = rho * vel[d]
    cells[4] = pr / (gamma - 1) + 0.5 * rho * sum(v * v for v in vel)
    om = oracle_uniform(mesh, mesh.from_cells(cells))
    out = om.courant_fine(3, O.make_params(gamma=gamma, courant_factor=C))
    c = np.sqrt(gamma * pr / rho)
    ctot = sum(abs(v) + c for v in vel)
    dt = 2 * C * mesh.dx / (ctot * (1 + np.sqrt(1 + 2 * C * 1e-4)))
    assert abs(out[4] - dt) <= 1e-14 * dt
    vol = mesh.dx ** 3
    ncell = 8 ** 3
    assert abs(out[0] - rho * vol * ncell) <= 1e-13 * rho * vol * ncell
    eint = pr / (gamma - 1) * vol * ncell
    assert abs(out[2] - eint) <= 1e-12 * eint


# ---------------------------------------------------------------------------
# Sod shock tube against the exact Riemann solution (Toro, ch. 4)
# ---------------------------------------------------------------------------
def exact_riemann(rl, ul, pl, rr, ur, pr, gamma, xi):
    cl, cr = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)

    def fk(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((gamma + 1) * rk), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)

    lo, hi = 1e-12, 100.0
    for _ in range(200):
        pm = 0.5 * (lo + hi)
        if fk(pm, rl, pl, cl) + fk(pm, rr, pr, cr) + ur - ul > 0:
            hi = pm
        else:
            lo = pm
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (fk(ps, rr, pr, cr) - fk(ps, rl, pl, cl))
    out = np.zeros((3, len(xi)))
    g = gamma
    for i, s in enumerate(xi):
        if s < us:  # left of contact
            if ps > pl:
                rs = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
                S = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                out[:, i] = (rl, ul, pl) if s < S else (rs, us, ps)
            else:
                rs = rl * (ps / pl) ** (1 / g)
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                if s < ul - cl:
                    out[:, i] = (rl, ul, pl)
                elif s > us - cs:
                    out[:, i] = (rs, us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    out[:, i] = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                rs = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
                S = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                out[:, i] = (rr, ur, pr) if s > S else (rs, us, ps)
            else:
                rs = rr * (ps / pr) ** (1 / g)
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                if s > ur + cr:
                    out[:, i] = (rr, ur, pr)
                elif s < us + cs:
                    out[:, i] = (rs, us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    out[:, i] = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
    return out


def test_sod_matches_exact_solution():
    """1024 cells, MinMod + HLLC, Courant 0.8, t=0.2: L1 density error < 2e-3."""
    from ramses2_amd.hydro import newdt
    mesh = UniformMesh(1, 10)
    p = HydroParams(gamma=1.4, courant_factor=0.8, slope_type=1, riemann="hllc")
    om = oracle_uniform(mesh, mesh.from_cells(ic.sod_1d(1024)), periodic=False)
    po = oracle_params(p)
    t, dtold = 0.0, 0.0
    while t < 0.2:
        c = om.courant_fine(10, po)
        dt = newdt(mesh.dx, p, c[4], t, 0.2, 3, dtold)
        om.set_unew(10)
        om.godunov_fine(10, dt, po)
        om.set_uold(10)
        t += dt
        dtold = dt
    assert abs(t - 0.2) < 1e-9
    cells = mesh.to_cells(om.uold)
    x = (np.arange(1024) + 0.5) / 1024
    ex = exact_riemann(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4, (x - 0.5) / t)
    l1 = np.mean(np.abs(cells[0] - ex[0]))
    assert l1 < 2e-3, l1
    # mass is conserved up to the boundary fluxes (zero-gradient ends carry none at rest)
    assert abs(cells[0].sum() - ic.sod_1d(1024)[0].sum()) < 1e-12 * 1024


def test_sub_box_equals_full_level():
    """A periodic computational box of 8 octs per dimension inside the 16-oct
    level (box_ckey_max = 8) steps exactly like the full 8-oct level with the
    same dx: the box bounds alone decide the periodic wrap."""
    periodic = True
    rng = np.random.default_rng(21)
    cells = ic.random_state((16, 16, 16), rng)
    p = HydroParams(gamma=1.4, courant_factor=0.5, slope_type=2, riemann="hllc")
    full = UniformMesh(3, 4)                       # 8 octs per dimension, dx = 1/16
    sub = UniformMesh(3, 5, boxlen=2.0, n=8)       # 8 of 16 octs, dx = 2/32
    assert full.dx == sub.dx and sub.box() == (5, [0, 0, 0], [8, 8, 8])
    out = []
    for m in (full, sub):
        om = oracle_uniform(m, m.from_cells(cells), periodic=periodic)
        om.godunov_fine(m.level, 1e-3, oracle_params(p))
        out.append(m.to_cells(om.unew))
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("threads", [1, 4])
def test_uniform_omp_baseline_equals_godunov_fine(threads):
    """The CPU baseline bench.py times (ora_godunov_uniform_omp: the same
    per-oct arithmetic, OpenMP over octs, neighbours from an oct map instead
    of the hash) gives godunov_fine's result bit for bit on a periodic level,
    with one thread and with several."""
    from ramses2_amd import UniformMesh, ic
    from tests.helpers import oracle_params, oracle_uniform
    mesh = UniformMesh(3, 4)
    p = HydroParams(gamma=5.0 / 3.0, courant_factor=0.5, slope_type=2, riemann="hllc")
    po = oracle_params(p)
    u0 = mesh.from_cells(ic.turbulence_3d(16, seed=3))
    a = oracle_uniform(mesh, u0.copy())
    b = oracle_uniform(mesh, u0.copy())
    dt = a.courant_fine(4, po)[4]
    a.set_unew(4)
    a.godunov_fine(4, dt, po)
    b.set_unew(4)
    b.godunov_uniform_omp(4, dt, po, (mesh.octmap + 1).astype(np.int32), threads)
    assert np.array_equal(a.unew, b.unew)


def _linear_u1(ndim, a0, g):
    """interpol_hydro's input for a field linear in each conserved variable:
    cell 0 the parent cell, cells 2d-1 and 2d its neighbours below and above
    in dimension d (one coarse cell away), value a0 + (offset) * g[d]."""
    u1 = np.zeros((2 * ndim + 1, 5))
    u1[0] = a0
    for d in range(ndim):
        u1[2 * d + 1] = a0 - g[d]
        u1[2 * d + 2] = a0 + g[d]
    return u1


def _children(ndim, a0, g, scale=1.0):
    """Children of the parent cell at offsets of a quarter coarse cell:
    a0 + scale * sum_d (+-1/4) g[d], in the cell order 1 + ix + 2 iy + 4 iz."""
    out = np.zeros((2 ** ndim, 5))
    for ind in range(2 ** ndim):
        s = [((ind >> d) & 1) - 0.5 for d in range(ndim)]
        out[ind] = a0 + scale * sum(0.5 * s[d] * g[d] for d in range(ndim))
    return out


@pytest.mark.parametrize("ndim", [1, 3])
def test_interpolation_of_linear_fields(ndim):
    """interpol_hydro (hydro/interpol_hydro.f90:5-113) on conservative
    variables (interpol_var=0): MinMod (1) and the unlimited central slope (3)
    reproduce a linear field exactly in the children; MonCen (2) does too
    when its corner limiter does not bite (the largest corner deviation,
    sum_d |g_d|/2, within the largest neighbour deviation max_d |g_d|), and
    scales every slope by max|g| / (sum|g|/2) when it does; type 0 injects
    the parent value."""
    rng = np.random.default_rng(7)
    a0 = np.array([1.3, 0.2, -0.1, 0.05, 2.5])
    for _ in range(5):
        g = [rng.uniform(-0.1, 0.1, 5) for _ in range(ndim)]
        for it in (1, 3):
            u2, _ = O.interpol_hydro(ndim, _linear_u1(ndim, a0, g), 0, it)
            assert np.allclose(u2, _children(ndim, a0, g), rtol=0, atol=1e-15), it
        u2, _ = O.interpol_hydro(ndim, _linear_u1(ndim, a0, g), 0, 0)
        assert np.array_equal(u2, np.tile(a0, (2 ** ndim, 1)))
    # MonCen along one axis (never limited) and along the diagonal (limited)
    g1 = [np.full(5, 0.08)] + [np.zeros(5)] * (ndim - 1)
    u2, _ = O.interpol_hydro(ndim, _linear_u1(ndim, a0, g1), 0, 2)
    assert np.allclose(u2, _children(ndim, a0, g1), rtol=0, atol=1e-15)
    gd = [np.full(5, 0.08)] * ndim
    u2, _ = O.interpol_hydro(ndim, _linear_u1(ndim, a0, gd), 0, 2)
    lim = min(1.0, 0.08 / (0.5 * ndim * 0.08))
    assert np.allclose(u2, _children(ndim, a0, gd, lim), rtol=0, atol=1e-15)
