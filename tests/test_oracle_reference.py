"""The CPU oracle against the reference's own kernels.

Golden vectors (tests/golden/ref_kernels.npz, written by make_golden.py from
the flang build of hydro/umuscl.f90, hydro/godunov_utils.f90 and
hydro/interpol_hydro.f90) must be reproduced bit for bit.  Where the reference
build is present (oracle/_ref, this container only) a seeded random sweep is
compared live as well.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_kernels.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("ndim", [1, 3])
def test_unsplit_golden(gold, ndim):
    U = gold["unsplit%dd_u" % ndim]
    G = gold["unsplit%dd_g" % ndim]
    P = gold["unsplit%dd_par" % ndim]
    F = gold["unsplit%dd_flux" % ndim]
    for i in range(U.shape[0]):
        dx, dt, difmag, st, rs = P[i]
        p = O.make_params(gamma=1.4, courant_factor=0.8, slope_type=int(st), riemann=int(rs), difmag=difmag)
        f = O.unsplit(ndim, U[i], G[i], dx, dt, p)
        assert np.array_equal(f, F[i]), (ndim, i)


@pytest.mark.parametrize("solver", [1, 2, 3])
def test_riemann_golden(gold, solver):
    ql, qr, F = gold["riemann_ql"], gold["riemann_qr"], gold["riemann%d_f" % solver]
    p = O.make_params(gamma=1.4, smax_cap=3.0)
    for i in range(ql.shape[0]):
        f = O.riemann(solver, ql[i], qr[i], p)
        assert np.array_equal(f, F[i]), (solver, i, f, F[i])


def test_cmpdt_golden(gold):
    p = O.make_params(gamma=1.4, courant_factor=0.5)
    for i in range(gold["cmpdt_uu"].shape[0]):
        dt, uu = O.cmpdt(3, gold["cmpdt_uu"][i], gold["cmpdt_gg"][i], 1.0 / 128, p)
        assert dt == gold["cmpdt_dt"][i]
        assert np.array_equal(uu, gold["cmpdt_uu_out"][i])


def test_interpol_golden(gold):
    for i in range(gold["interp_u1"].shape[0]):
        it, iv = i % 4, (i // 4) % 2
        u2, _ = O.interpol_hydro(3, gold["interp_u1"][i], iv, it)
        assert np.array_equal(u2, gold["interp_u2"][i]), (i, it, iv)


def _states(rng, shape, near_vacuum=False):
    rho = np.exp(rng.normal(0, 1, shape))
    vel = rng.normal(0, 2, shape + (3,))
    pr = np.exp(rng.normal(0, 1, shape))
    if near_vacuum:
        rho[rng.random(shape) < 0.2] = 1e-12
        pr[rng.random(shape) < 0.2] = 1e-14
    u = np.zeros(shape + (5,), order="F")
    u[..., 0] = rho
    u[..., 1:4] = rho[..., None] * vel
    u[..., 4] = pr / 0.4 + 0.5 * rho * (vel ** 2).sum(-1)
    return u


@pytest.mark.parametrize("ndim", [1, 3])
def test_unsplit_live_reference(ndim):
    R = O.ref_lib(ndim)
    if R is None:
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(ndim)
    shp = O.stencil_shape(ndim)
    for trial in range(60):
        for st in (1, 2, 3):
            for rs in (1, 2, 3):
                u = _states(rng, shp, near_vacuum=(trial % 7 == 0))
                g = np.zeros(shp + (ndim,), order="F")
                if trial % 5 == 1:
                    g[...] = rng.normal(0, 1, ndim)
                dx, dt = 1 / 64, 1e-3 * (1 + rng.random())
                difmag = 0.1 if trial % 11 == 3 else 0.0
                p = O.make_params(gamma=1.4, courant_factor=0.8, slope_type=st, riemann=rs, difmag=difmag)
                f1 = O.unsplit(ndim, u, g, dx, dt, p)
                f2 = np.zeros_like(f1)
                R.ref_unsplit(2, O._ptr(u, O._dp), O._ptr(g, O._dp), O._ptr(f2, O._dp), dx, dt, 1.4, 1e-10,
                              1e-10, 0.8, difmag, st, rs)
                assert np.array_equal(f1, f2), (trial, st, rs)


def test_interpol_live_reference_1d():
    R = O.ref_lib(1)
    if R is None:
        pytest.skip("reference build (oracle/_ref) not present")
    rng = np.random.default_rng(9)
    for i in range(40):
        u1 = np.asfortranarray(_states(rng, (3,)))
        it, iv = i % 4, (i // 4) % 2
        a = np.asfortranarray(u1.copy())
        b = np.zeros((2, 5), order="F")
        R.ref_interpol_hydro(O._ptr(a, O._dp), O._ptr(b, O._dp), iv, it, 1e-10)
        u2, _ = O.interpol_hydro(1, u1, iv, it)
        assert np.array_equal(u2, b)
