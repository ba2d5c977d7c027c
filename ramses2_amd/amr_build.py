"""Static AMR meshes refined with the reference's gradient criterion.

A mesh is built level by level from an analytic state: a cell of level l is
flagged when hydro_refine's error (hydro/godunov_utils.f90:6-227, applied per
direction by hydro_flag, hydro/hydro_flag.f90:74-150) for density or pressure
exceeds err_grad_d / err_grad_p, with the face neighbours sampled at the same
resolution.  Flags are dilated by nexpand cells, then limited to cells whose
oct has all 26 neighbour octs on the level, so that every oct of level l+1
finds the parent cells (and their neighbours) its ghost octs interpolate
from.  Flagged cells of level l become the octs of level l+1.

The state on the mesh is the analytic state at the cell centres, with every
refined cell set to the average of its child oct (upload_fine).
"""
import numpy as np

from .mesh import AmrMesh, prim_to_cons

OFF8 = np.array([[(i & 1), (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], dtype=np.int64)


def refine_error(ug, um, ud, floor):
    """2 max(|(d-m)/(d+m+floor)|, |(m-g)/(m+g+floor)|), hydro_refine's error
    for one variable along one direction (godunov_utils.f90:122-136)."""
    return 2.0 * np.maximum(np.abs((ud - um) / (ud + um + floor)), np.abs((um - ug) / (um + ug + floor)))


def sedov_profile(x, y, z, radius=0.3, gamma=1.4, rho0=1.0, p0=1e-5, e_blast=1.0, width=0.0, centre=0.5):
    """Primitive state (rho, vx, vy, vz, p) of a Sedov-Taylor blast of the
    given shock radius in the periodic unit box: post-shock values of a strong
    shock moving at u_s = (2/5) R/t with R = 1.15 (E t^2/rho0)^(1/5)
    (gamma = 1.4), an interior density ~ (r/R)^3, linear velocity and a
    pressure plateau; a sharp shock unless width > 0 (tanh joint)."""
    dxc = x - centre
    dyc = y - centre
    dzc = z - centre
    r = np.sqrt(dxc * dxc + dyc * dyc + dzc * dzc)
    t = (radius / 1.15) ** 2.5 * np.sqrt(rho0 / e_blast)
    us = 0.4 * radius / t
    rho_s = (gamma + 1.0) / (gamma - 1.0) * rho0
    p_s = 2.0 / (gamma + 1.0) * rho0 * us * us
    v_s = 2.0 / (gamma + 1.0) * us
    if width:
        inside = 0.5 * (1.0 - np.tanh((r - radius) / width))
    else:
        inside = (r < radius).astype(np.float64)
    s = np.minimum(r / radius, 1.0)
    rho = inside * np.maximum(rho_s * s ** 3, 1e-3 * rho0) + (1.0 - inside) * rho0
    p = inside * p_s * (0.36 + 0.64 * s ** 3) + (1.0 - inside) * p0
    vr = inside * v_s * s
    rr = np.where(r > 0, r, 1.0)
    return rho, vr * dxc / rr, vr * dyc / rr, vr * dzc / rr, p


def _shift(a, k, axis):
    return np.roll(a, k, axis=axis)


def _dilate(f, n):
    for _ in range(n):
        for ax in range(3):
            f = f | _shift(f, 1, ax) | _shift(f, -1, ax)
    return f


def _erode(f):
    for ax in range(3):
        f = f & _shift(f, 1, ax) & _shift(f, -1, ax)
    return f


def build_levels(levelmin, levelmax, profile, err_grad_d=0.1, err_grad_p=0.1, nexpand=1, floor_d=1e-10,
                 floor_p=1e-10):
    """Oct keys {level: (n,3) int array} of a periodic unit-box mesh: level
    levelmin complete, finer levels where profile(x, y, z) is flagged."""
    n0 = 2 ** (levelmin - 1)
    ii = np.arange(n0)
    keys = {levelmin: np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)}
    for l in range(levelmin, levelmax):
        k = keys[l]
        if len(k) == 0:
            break
        nc = 2 ** l                      # cells per dimension at level l
        dx = 1.0 / nc
        cells = (2 * k[:, None, :] + OFF8[None, :, :]).reshape(-1, 3)
        xc = (cells + 0.5) * dx
        ok = np.zeros(len(cells), dtype=bool)
        mid = profile(xc[:, 0], xc[:, 1], xc[:, 2])
        for d in range(3):
            e = np.zeros(3)
            e[d] = dx
            lo = profile(xc[:, 0] - e[0], xc[:, 1] - e[1], xc[:, 2] - e[2])
            hi = profile(xc[:, 0] + e[0], xc[:, 1] + e[1], xc[:, 2] + e[2])
            if err_grad_d >= 0:
                ok |= refine_error(lo[0], mid[0], hi[0], floor_d) > err_grad_d
            if err_grad_p >= 0:
                ok |= refine_error(lo[4], mid[4], hi[4], floor_p) > err_grad_p
        flag = np.zeros((nc, nc, nc), dtype=bool)          # [x, y, z]
        c = cells[ok]
        flag[c[:, 0], c[:, 1], c[:, 2]] = True
        flag = _dilate(flag, nexpand)
        octs = np.zeros((nc // 2,) * 3, dtype=bool)
        octs[k[:, 0], k[:, 1], k[:, 2]] = True
        allowed = _erode(octs)
        allowed = allowed.repeat(2, 0).repeat(2, 1).repeat(2, 2)
        flag &= allowed
        keys[l + 1] = np.argwhere(flag).astype(np.int64)
        if len(keys[l + 1]) == 0:
            del keys[l + 1]
            break
    return keys


def amr_state(am: AmrMesh, profile, gamma=1.4):
    """uold (ngrid, 5, 8) of the profile at the cell centres, refined cells
    restricted from their children, finest level first."""
    u = np.zeros((am.ngrid, 5, 8))
    for l in am.levels:
        x = am.cell_centres(l)                                   # (noct, 8, 3)
        rho, vx, vy, vz, p = profile(x[..., 0], x[..., 1], x[..., 2])
        c = prim_to_cons(rho, (vx, vy, vz), p, gamma)             # (5, noct, 8)
        i0 = am.head[l] - 1
        u[i0:i0 + am.noct[l]] = np.transpose(c, (1, 0, 2))
    for l in reversed(am.levels[1:]):
        i0 = am.head[l] - 1
        kf = am.ckey[i0:i0 + am.noct[l]].astype(np.int64)
        par, ind = am.parent_cells(l, kf)
        avg = u[i0:i0 + am.noct[l]].mean(axis=2)                  # (noct, 5)
        u[par, :, ind] = avg
    return u


def sedov_amr(levelmin=6, levelmax=10, radius=0.3, gamma=1.4, err_grad=0.1):
    """The AMR Sedov workload: mesh and initial state."""
    prof = lambda x, y, z: sedov_profile(x, y, z, radius=radius, gamma=gamma)  # noqa: E731
    keys = build_levels(levelmin, levelmax, prof, err_grad, err_grad)
    am = AmrMesh(3, keys)
    return am, amr_state(am, prof, gamma)
