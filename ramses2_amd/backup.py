"""RAMSES II backup files: exchange meshes and hydro states with the reference.

The byte layout restates the reference's writers and readers:
- ``amr.NNNNN`` (backup_amr, amr/output_amr.f90:666-735): Fortran stream
  (no record markers), int32 ndim, levelmin, nlevelmax, then noct(l) for
  l = levelmin..nlevelmax, then per oct, level by level, int32 ckey(1:ndim)
  and an int32 whose bit ind-1 is set when cell ind is refined;
- ``hydro.NNNNN`` (backup_hydro, hydro/output_hydro.f90:167-246): int32
  ndim, nvar, levelmin, nlevelmax, noct(l), then per oct uold(1:twotondim,
  1:nvar) as float64, cell index fastest;
- the restart reader seeks oct i at byte 12 + 4 nlev + (4 ndim + 4)(i-1) of
  the amr file and 16 + 4 nlev + 8 twotondim nvar (i-1) of the hydro file
  (amr/init_refine_restart.f90:386-441; its POS values are 1-based);
- ``info.txt`` (output_info, amr/output_amr.f90:790-838): the run's
  parameters as formatted text; the reader takes time, boxlen, gamma,
  levelmin, levelmax from it.
File numbers are title(n): five digits, zero padded (amr/title.f90).
One file per rank; this module writes and reads one (ncpu = 1).
"""
import os
import struct

import numpy as np

from .mesh import NVAR


def title(n):
    """title(n) of amr/title.f90: n as five digits, zero padded."""
    return "%05d" % int(n)


def write_amr(path, ndim, levelmin, nlevelmax, levels):
    """levels: {l: (ckey (n, ndim) int, refined (n, 2^ndim) bool)} for
    l = levelmin..nlevelmax (missing levels have no octs)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<3i", ndim, levelmin, nlevelmax))
        for l in range(levelmin, nlevelmax + 1):
            f.write(struct.pack("<i", len(levels[l][0]) if l in levels else 0))
        for l in range(levelmin, nlevelmax + 1):
            if l not in levels or len(levels[l][0]) == 0:
                continue
            ck, ref = levels[l]
            n = len(ck)
            rec = np.zeros((n, ndim + 1), dtype="<i4")
            rec[:, :ndim] = np.asarray(ck)[:, :ndim]
            bits = np.asarray(ref, dtype=np.int64)[:, :2 ** ndim] << np.arange(2 ** ndim)
            rec[:, ndim] = bits.sum(axis=1)
            f.write(rec.tobytes())


def read_amr(path):
    """-> ndim, levelmin, nlevelmax, {l: (ckey, refined)}."""
    with open(path, "rb") as f:
        data = f.read()
    ndim, levelmin, nlevelmax = struct.unpack_from("<3i", data, 0)
    nlev = nlevelmax - levelmin + 1
    noct = struct.unpack_from("<%di" % nlev, data, 12)
    off = 12 + 4 * nlev
    levels = {}
    for k, l in enumerate(range(levelmin, nlevelmax + 1)):
        n = noct[k]
        if n == 0:
            continue
        rec = np.frombuffer(data, dtype="<i4", count=n * (ndim + 1), offset=off).reshape(n, ndim + 1)
        off += 4 * (ndim + 1) * n
        ref = ((rec[:, ndim:ndim + 1] >> np.arange(2 ** ndim)) & 1).astype(bool)
        levels[l] = (rec[:, :ndim].astype(np.int32), ref)
    if off != len(data):
        raise ValueError("%s: %d bytes, expected %d" % (path, len(data), off))
    return ndim, levelmin, nlevelmax, levels


def write_hydro(path, ndim, levelmin, nlevelmax, levels_u, nvar=NVAR):
    """levels_u: {l: uold of the level's octs, (n, nvar, 2^ndim) float64}."""
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", ndim, nvar, levelmin, nlevelmax))
        for l in range(levelmin, nlevelmax + 1):
            f.write(struct.pack("<i", len(levels_u[l]) if l in levels_u else 0))
        for l in range(levelmin, nlevelmax + 1):
            if l in levels_u and len(levels_u[l]):
                u = np.ascontiguousarray(levels_u[l], dtype="<f8")
                assert u.shape[1:] == (nvar, 2 ** ndim)
                f.write(u.tobytes())


def read_hydro(path):
    """-> ndim, nvar, levelmin, nlevelmax, {l: uold (n, nvar, 2^ndim)}."""
    with open(path, "rb") as f:
        data = f.read()
    ndim, nvar, levelmin, nlevelmax = struct.unpack_from("<4i", data, 0)
    nlev = nlevelmax - levelmin + 1
    noct = struct.unpack_from("<%di" % nlev, data, 16)
    off = 16 + 4 * nlev
    ttd = 2 ** ndim
    levels = {}
    for k, l in enumerate(range(levelmin, nlevelmax + 1)):
        n = noct[k]
        if n == 0:
            continue
        levels[l] = np.frombuffer(data, dtype="<f8", count=n * nvar * ttd, offset=off).reshape(n, nvar, ttd).copy()
        off += 8 * nvar * ttd * n
    if off != len(data):
        raise ValueError("%s: %d bytes, expected %d" % (path, len(data), off))
    return ndim, nvar, levelmin, nlevelmax, levels


def _e23(x):
    """Fortran E23.15: 0.ddddddddddddddde+xx, right-aligned in 23 columns."""
    if x == 0:
        m, e = 0.0, 0
    else:
        e = int(np.floor(np.log10(abs(x)))) + 1
        m = x / 10.0 ** e
        if abs(m) >= 1.0:            # rounding up to 1.0
            m /= 10.0
            e += 1
    s = "%.15f" % abs(m)
    s = ("-" if x < 0 else "") + s + ("E%+03d" % e)
    return s.rjust(23)


def write_info(path, ndim, levelmin, levelmax, ngridmax, nstep_coarse, boxlen, time, gamma, ncpu=1, nfile=1):
    """info.txt in the layout of output_info (amr/output_amr.f90:790-838);
    cosmology fields are those of a non-cosmological run."""
    lines = [
        "nfile       =%11d" % nfile, "ncpu        =%11d" % ncpu, "ndim        =%11d" % ndim,
        "levelmin    =%11d" % levelmin, "levelmax    =%11d" % levelmax, "ngridmax    =%11d" % ngridmax,
        "nstep_coarse=%11d" % nstep_coarse, "",
        "boxlen      =" + _e23(boxlen), "time        =" + _e23(time), "texp        =" + _e23(0.0),
        "aexp        =" + _e23(1.0), "H0          =" + _e23(1.0), "omega_m     =" + _e23(1.0),
        "omega_l     =" + _e23(0.0), "omega_k     =" + _e23(0.0), "omega_b     =" + _e23(0.0),
        "gamma       =" + _e23(gamma), "unit_l      =" + _e23(1.0), "unit_d      =" + _e23(1.0),
        "unit_t      =" + _e23(1.0), "",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_info(path):
    out = {}
    with open(path) as f:
        for line in f:
            if "=" not in line:
                continue
            k, v = line.split("=", 1)
            k, v = k.strip(), v.strip()
            try:
                out[k] = int(v)
            except ValueError:
                out[k] = float(v.replace("D", "E"))
    return out


def write_backup(root, nout, amesh, uold, time, gamma, nstep_coarse=0, ngridmax=None):
    """backup_NNNNN/{amr,hydro}.00001 and info.txt of an AmrMesh and its uold."""
    d = os.path.join(root, "backup_" + title(nout))
    os.makedirs(d, exist_ok=True)
    lv, lu = {}, {}
    for l in amesh.levels:
        i0 = amesh.head[l] - 1
        sl = slice(i0, i0 + amesh.noct[l])
        lv[l] = (amesh.ckey[sl], amesh.refined[sl].astype(bool))
        lu[l] = uold[sl]
    write_amr(os.path.join(d, "amr." + title(1)), 3, amesh.levelmin, amesh.levelmax, lv)
    write_hydro(os.path.join(d, "hydro." + title(1)), 3, amesh.levelmin, amesh.levelmax, lu)
    write_info(os.path.join(d, "info.txt"), 3, amesh.levelmin, amesh.levelmax, ngridmax or amesh.ngrid,
               nstep_coarse, amesh.boxlen, time, gamma)
    return d


def read_backup(root, nout):
    """-> (AmrMesh, uold, info) from backup_NNNNN (one file per kind).  The
    octs of each level are re-ordered by the mesh (Morton order); uold follows."""
    from .mesh import AmrMesh
    d = os.path.join(root, "backup_" + title(nout))
    ndim, lmin, lmax, lv = read_amr(os.path.join(d, "amr." + title(1)))
    ndim2, nvar, lmin2, lmax2, lu = read_hydro(os.path.join(d, "hydro." + title(1)))
    if (ndim, lmin, lmax) != (ndim2, lmin2, lmax2) or nvar != NVAR or ndim != 3:
        raise ValueError("amr and hydro files disagree, or not a 3D hydro run with nvar=5")
    info = read_info(os.path.join(d, "info.txt"))
    am = AmrMesh(3, {l: lv[l][0] for l in lv}, boxlen=info.get("boxlen", 1.0))
    u = np.zeros((am.ngrid, NVAR, 8))
    for l in lv:
        idx = am.find(l, lv[l][0])
        u[idx] = lu[l]
        if not np.array_equal(am.refined[idx].astype(bool), lv[l][1]):
            raise ValueError("refined flags of level %d do not match the octs present" % l)
    return am, u, info
