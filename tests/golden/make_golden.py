"""Write tests/golden/ref_kernels.npz: input/output vectors of the reference's
own hydro kernels (hydro/umuscl.f90 unsplit, hydro/godunov_utils.f90
riemann_llf/hll/hllc and cmpdt, hydro/interpol_hydro.f90 interpol_hydro),
computed by the flang build in oracle/_ref (make -C oracle ref, which needs
the reference tree).  The file holds data only: seeded inputs and the
reference's outputs.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_kernels.npz")


def states(rng, shape, near_vacuum=False):
    rho = np.exp(rng.normal(0, 1, shape))
    vel = rng.normal(0, 2, shape + (3,))
    pr = np.exp(rng.normal(0, 1, shape))
    if near_vacuum:
        rho[rng.random(shape) < 0.2] = 1e-12
        pr[rng.random(shape) < 0.2] = 1e-14
    u = np.zeros(shape + (5,))
    u[..., 0] = rho
    u[..., 1:4] = rho[..., None] * vel
    u[..., 4] = pr / 0.4 + 0.5 * rho * (vel ** 2).sum(-1)
    return u


def main():
    rng = np.random.default_rng(20261202)
    data = {}
    for ndim in (1, 3):
        R = O.ref_lib(ndim)
        if R is None:
            raise SystemExit("oracle/_ref not built: run make -C oracle ref")
        shp = O.stencil_shape(ndim)
        cases = []
        for case in range(12):
            st = 1 + case % 3
            rs = 1 + (case // 3) % 3
            u = np.asfortranarray(states(rng, shp, near_vacuum=(case % 4 == 3)))
            g = np.zeros(shp + (ndim,), order="F")
            if case % 5 == 1:
                g[...] = rng.normal(0, 1, ndim)
            dx, dt = 1.0 / 64, 1e-3 * (1 + rng.random())
            difmag = 0.1 if case % 6 == 5 else 0.0
            flux = np.zeros(O.flux_shape(ndim) + (5, ndim), order="F")
            R.ref_unsplit(2, O._ptr(u, O._dp), O._ptr(g, O._dp), O._ptr(flux, O._dp), dx, dt, 1.4, 1e-10,
                          1e-10, 0.8, difmag, st, rs)
            cases.append((u, g, np.array([dx, dt, difmag, st, rs]), flux))
        data["unsplit%dd_u" % ndim] = np.stack([c[0] for c in cases])
        data["unsplit%dd_g" % ndim] = np.stack([c[1] for c in cases])
        data["unsplit%dd_par" % ndim] = np.stack([c[2] for c in cases])
        data["unsplit%dd_flux" % ndim] = np.stack([c[3] for c in cases])
    R = O.ref_lib(3)
    # Riemann solvers
    n = 200
    ql = np.zeros((n, 5))
    qr = np.zeros((n, 5))
    for q in (ql, qr):
        q[:, 0] = np.exp(rng.normal(0, 1, n))
        q[:, 1:4] = rng.normal(0, 2, (n, 3))
        q[:, 4] = np.exp(rng.normal(0, 1, n))
    ql[::17, 0] = 1e-13
    qr[::13, 4] = 1e-16
    for solver in (1, 2, 3):
        f = np.zeros((n, 6))
        for i in range(n):
            a, b = ql[i].copy(), qr[i].copy()
            o = np.zeros(6)
            R.ref_riemann(solver, O._ptr(a, O._dp), O._ptr(b, O._dp), O._ptr(o, O._dp), 1.4, 1e-10, 1e-10,
                          3.0)
            f[i] = o
        data["riemann%d_f" % solver] = f
    data["riemann_ql"] = ql
    data["riemann_qr"] = qr
    # cmpdt
    uu = states(rng, (n,))
    gg = rng.normal(0, 1, (n, 3))
    gg[::2] = 0.0
    dts = np.zeros(n)
    uus = np.zeros((n, 5))
    for i in range(n):
        a = uu[i].copy()
        g = gg[i].copy()
        d = np.zeros(1)
        R.ref_cmpdt(O._ptr(a, O._dp), O._ptr(g, O._dp), 1.0 / 128, 1.4, 1e-10, 1e-10, 0.5, O._ptr(d, O._dp))
        dts[i] = d[0]
        uus[i] = a
    data["cmpdt_uu"] = uu
    data["cmpdt_gg"] = gg
    data["cmpdt_dt"] = dts
    data["cmpdt_uu_out"] = uus
    # interpol_hydro (3D)
    m = 60
    u1 = states(rng, (m, 7))
    u2s = []
    for i in range(m):
        it = i % 4
        iv = (i // 4) % 2
        a = np.asfortranarray(u1[i].copy())
        b = np.zeros((8, 5), order="F")
        R.ref_interpol_hydro(O._ptr(a, O._dp), O._ptr(b, O._dp), iv, it, 1e-10)
        u2s.append(b)
    data["interp_u1"] = u1
    data["interp_u2"] = np.stack(u2s)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
