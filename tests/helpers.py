"""Shared helpers of the test suite (oracle meshes, error metrics)."""
import numpy as np

from oracle import oracle as O
from ramses2_amd.mesh import UniformMesh


def oracle_uniform(mesh: UniformMesh, uold, periodic=True, bound_type=2, unew=None):
    """An oracle mesh holding one uniform level (levelmin = levelmax = mesh.level)."""
    nl = mesh.level
    head = [0] * nl
    noct = [0] * nl
    head[nl - 1] = 1
    noct[nl - 1] = mesh.noct
    per = (1, 1, 1) if periodic else (0, 0, 0)
    bt = ((bound_type, bound_type),) * 3
    return O.OracleMesh(mesh.ndim, mesh.ckey, np.full(mesh.noct, nl), np.zeros((mesh.noct, 2 ** mesh.ndim)),
                        uold, head, noct, levelmin=nl, boxlen=mesh.boxlen, periodic=per, bound_type=bt,
                        unew=unew, box=mesh.box())


def oracle_params(p):
    return O.make_params(gamma=p.gamma, smallr=p.smallr, smallc=p.smallc, courant_factor=p.courant_factor,
                         slope_type=p.slope_type, riemann=p.riemann_id, difmag=p.difmag)


def rel_err_per_var(a, b):
    """max_cells |a-b| / max_cells |b| for each conserved variable (absolute when b is 0)."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for v in range(a.shape[1]):
        d = np.max(np.abs(a[:, v] - b[:, v]))
        s = np.max(np.abs(b[:, v]))
        out.append(d / s if s > 0 else d)
    return np.array(out)


def rel_err_per_cell(a, b, floor=1e-6):
    """max over cells of |a-b| / max(|b|, floor * max_cells |b|) for each
    conserved variable: the relative error of every cell, with a floor that
    keeps components crossing zero (momenta) from dividing by ~0."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for v in range(a.shape[1]):
        s = np.max(np.abs(b[:, v]))
        den = np.maximum(np.abs(b[:, v]), floor * s if s > 0 else 1.0)
        out.append(np.max(np.abs(a[:, v] - b[:, v]) / den))
    return np.array(out)


def l1_rel_per_var(a, b):
    """sum |a-b| / sum |b| per conserved variable."""
    out = []
    for v in range(a.shape[1]):
        s = np.sum(np.abs(b[:, v]))
        d = np.sum(np.abs(a[:, v] - b[:, v]))
        out.append(d / s if s > 0 else d)
    return np.array(out)
