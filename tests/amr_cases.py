"""Two-level AMR test meshes and the oracle's amr_step (nsubcycle=1)."""
import numpy as np

from ramses2_amd.mesh import AmrMesh, prim_to_cons
from tests.helpers import oracle_params

from oracle import oracle as O


def two_level_mesh(lo=4, region=(4, 12), offset=(0, 0, 0)):
    """Complete level lo plus a refined block of level lo+1 octs with keys in
    region^3 (shifted by offset, periodic)."""
    n = 2 ** (lo - 1)
    ii = np.arange(n)
    k4 = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    jj = np.arange(region[0], region[1])
    k5 = np.stack(np.meshgrid(jj, jj, jj, indexing="ij"), -1).reshape(-1, 3)
    k5 = (k5 + np.asarray(offset)) % (2 * n)
    return AmrMesh(3, {lo: k4, lo + 1: k5})


def three_level_mesh(lo=4):
    """Complete level lo, a block of level lo+1 octs (keys [4,12)^3) and inside
    it a block of level lo+2 octs (keys [12,20)^3): ghost octs of both fine
    levels have their parent cells and the parents' neighbours."""
    n = 2 ** (lo - 1)
    ii = np.arange(n)
    k4 = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    jj = np.arange(4, 12)
    k5 = np.stack(np.meshgrid(jj, jj, jj, indexing="ij"), -1).reshape(-1, 3)
    kk = np.arange(12, 20)
    k6 = np.stack(np.meshgrid(kk, kk, kk, indexing="ij"), -1).reshape(-1, 3)
    return AmrMesh(3, {lo: k4, lo + 1: k5, lo + 2: k6})


def low_face_mesh(shift=0, lo=4):
    """A box of 4 x 8 x 8 octs at level lo, non-periodic in x, whose x keys
    start at `shift` (the reference's non-periodic boxes start at a positive
    key, e.g. box_xmin=1); a refined block of level lo+1 touching the low x
    face: x keys [2 shift, 2 shift + 4), y and z keys [4, 12).  Its ghost octs
    next to that face take neighbouring parent cells from the coarse level's
    boundary octs."""
    k4 = np.stack(np.meshgrid(np.arange(shift, shift + 4), np.arange(8), np.arange(8), indexing="ij"),
                  -1).reshape(-1, 3)
    jj = np.arange(4, 12)
    k5 = np.stack(np.meshgrid(np.arange(2 * shift, 2 * shift + 4), jj, jj, indexing="ij"), -1).reshape(-1, 3)
    return AmrMesh(3, {lo: k4, lo + 1: k5}, box=(lo, [shift, 0, 0], [shift + 4, 8, 8]))


def smooth_state(am, seed, gamma=1.4, amp=0.3, noise=0.02, blast=False, shift_x=0.0):
    rng = np.random.default_rng(seed)
    u = np.zeros((am.ngrid, 5, 8))
    for l in am.levels:
        x = am.cell_centres(l)
        x[..., 0] = (x[..., 0] - shift_x) % am.boxlen
        i0 = am.head[l] - 1
        tp = 2 * np.pi
        rho = 1 + amp * np.sin(tp * x[..., 0]) * np.cos(tp * x[..., 1]) + noise * rng.random(x.shape[:2])
        vx = 0.5 * np.sin(tp * x[..., 1] + 1.0)
        vy = 0.4 * np.cos(tp * x[..., 2])
        vz = 0.3 * np.sin(tp * (x[..., 0] + x[..., 2]))
        p = 1 + 0.2 * np.cos(tp * x[..., 2]) + noise * rng.random(x.shape[:2])
        if blast:
            r2 = ((x - 0.5) ** 2).sum(-1)
            p = p + 20.0 * (r2 < 0.1 ** 2)
        c = prim_to_cons(rho, (vx, vy, vz), p, gamma)  # (5, noct, 8)
        u[i0:i0 + am.noct[l]] = np.transpose(c, (1, 0, 2))
    return u


def oracle_mesh(am, u, interpol_var=0, interpol_type=1, periodic=(1, 1, 1), bound_type=((2, 2), (2, 2), (2, 2))):
    head, noct = am.heads()
    return O.OracleMesh(3, am.ckey, am.lev, am.refined, u, head, noct, levelmin=am.levelmin,
                        boxlen=am.boxlen, interpol_var=interpol_var, interpol_type=interpol_type,
                        periodic=periodic, bound_type=bound_type, box=am.box)


def key_index(am):
    """{(level, key tuple): 0-based oct index} of a mesh."""
    return {(int(am.lev[i]), tuple(int(x) for x in am.ckey[i])): i for i in range(am.ngrid)}


def restrict_all(am, om):
    for l in reversed(am.levels[:-1]):
        om.upload_fine(l)


def oracle_step(am, om, p, dt=None):
    po = oracle_params(p)
    dts = []
    for l in am.levels:
        c = om.courant_fine(l, po)
        dts.append(min(am.dx(l) / p.smallc, c[4]))
        om.set_unew(l)
    if dt is None:
        dt = min(dts)
    for l in reversed(am.levels):
        om.godunov_fine(l, dt, po)
        om.set_uold(l)
        if l < am.levelmax:
            om.upload_fine(l)
    return dt


def leaf_total(am, u, var):
    """sum over leaf cells of u*vol."""
    s = 0.0
    for l in am.levels:
        i0 = am.head[l] - 1
        sl = slice(i0, i0 + am.noct[l])
        leaf = am.refined[sl] == 0
        s += (u[sl, var, :] * leaf).sum() * am.dx(l) ** 3
    return s
