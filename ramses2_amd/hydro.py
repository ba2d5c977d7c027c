"""Host side of the drop-in boundary: the reference's level-kernel call sequence
driven through the C-ABI of include/ramses2_hip.h.

``HipHydro`` mirrors the Fortran runner (metal/metal_runner.f90): one method
per rh_* entry point, same argument meaning, errors raised as exceptions (the
runner's ``error stop``).  ``UniformLevel`` drives a single-level mesh through
the reference's order of calls per time step (amr/amr_step.f90:276-455):
newdt_fine -> courant_fine, set_unew, godunov_fine, set_uold.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import call, ptr
from .mesh import NVAR, UniformMesh

RIEMANN_IDS = {"llf": 1, "hll": 2, "hllc": 3}


@dataclass
class HydroParams:
    """&HYDRO_PARAMS subset (amr/read_params.f90:318-354)."""

    gamma: float = 1.4
    courant_factor: float = 0.5
    smallr: float = 1e-10
    smallc: float = 1e-10
    slope_type: int = 1
    riemann: str = "llf"
    difmag: float = 0.0   # artificial diffusion (3D; rh_set_difmag)

    @property
    def riemann_id(self):
        return RIEMANN_IDS[self.riemann] if isinstance(self.riemann, str) else int(self.riemann)


class HipHydro:
    """The device-resident hydro state of this process (one per process, as one
    MPI rank owns one device in the reference)."""

    def __init__(self, ndim, ngridmax, ncachemax=0, hash_size=None):
        self.ndim = ndim
        self.ttd = 2 ** ndim
        self.ngridmax = ngridmax
        self.ncachemax = ncachemax
        self.hash_size = hash_size or 2 * (ngridmax + ncachemax)
        call("rh_init")
        call("rh_alloc_amr", ngridmax, ncachemax, NVAR, self.ttd, self.hash_size)
        self.box = None

    # ---- mesh set-up ----
    def upload_level_params(self, ckey_max, key_off, box_min, box_max, periodic):
        self.ckey_max = np.ascontiguousarray(ckey_max, dtype=np.int32)
        self.key_off = np.ascontiguousarray(key_off, dtype=np.int64)
        self.box_min = np.ascontiguousarray(box_min, dtype=np.int32)
        self.box_max = np.ascontiguousarray(box_max, dtype=np.int32)
        self.periodic = np.ascontiguousarray(periodic, dtype=np.int32)
        nlevelmax = len(self.ckey_max) - 1
        call("rh_upload_level_params", ptr(self.ckey_max, _lib._ip), ptr(self.key_off, _lib._lp),
             ptr(self.box_min, _lib._ip), ptr(self.box_max, _lib._ip), ptr(self.periodic, _lib._ip), nlevelmax)

    def set_boundaries(self, bound_type):
        bt = np.ascontiguousarray(np.asarray(bound_type, dtype=np.int32).reshape(6))
        call("rh_set_boundaries", ptr(bt, _lib._ip))

    def set_grid_device(self, uold, unew, grid):
        uold = np.ascontiguousarray(uold, dtype=np.float64)
        unew = np.ascontiguousarray(unew, dtype=np.float64)
        n = uold.shape[0]
        assert uold.shape == (n, NVAR, self.ttd) and unew.shape == uold.shape and len(grid) == n
        call("rh_set_grid_device", ptr(uold), ptr(unew), None, grid.ctypes.data_as(ctypes.c_void_p), n,
             NVAR, self.ttd)
        self.ngrid = n

    def transfer_grid_host(self, n=None):
        n = self.ngrid if n is None else n
        u = np.zeros((n, NVAR, self.ttd))
        call("rh_transfer_grid_host", ptr(u), None, n, NVAR, self.ttd)
        return u

    def transfer_unew_host(self, n=None):
        n = self.ngrid if n is None else n
        u = np.zeros((n, NVAR, self.ttd))
        call("rh_transfer_unew_host", ptr(u), n, NVAR, self.ttd)
        return u

    def insert_hash_all(self, head, num):
        call("rh_insert_hash_all", head, num, self.hash_size)

    def build_nbor(self, head, num, ilevel):
        l = ilevel - 1
        bmin = np.ascontiguousarray(self.box_min[l])
        bmax = np.ascontiguousarray(self.box_max[l])
        call("rh_build_nbor", head, num, self.hash_size, int(self.ckey_max[l]), int(self.key_off[l]),
             ptr(bmin, _lib._ip), ptr(bmax, _lib._ip), ptr(self.periodic, _lib._ip))

    def hydro_flag(self, head, num, p: HydroParams, err_grad_d, err_grad_p, floor_d=1e-10, floor_p=1e-10,
                   reset=True):
        """Refinement flags of octs head..head+num-1 from the device state
        (rh_hydro_flag, flag1 cleared first when reset): returns flag1
        (num, 8) int32 and the number of flagged cells."""
        if reset:
            call("rh_reset_flag1", head, num)
        call("rh_hydro_flag", head, num, p.gamma, p.smallr, p.smallc ** 2, float(err_grad_d), float(err_grad_p),
             float(floor_d), float(floor_p))
        n = ctypes.c_int32()
        call("rh_count_flag1", head, num, ctypes.byref(n))
        f = np.zeros((self.ngrid, 8), dtype=np.int32)
        call("rh_transfer_flag1_host", ptr(f, _lib._ip), self.ngrid)
        return f[head - 1:head - 1 + num], n.value

    def level_table(self, ilevel, which):
        """One of the level's device index tables (rh_get_level_table) as a
        flat int32 array."""
        n = ctypes.c_int64()
        L = _lib.load()
        L.rh_get_level_table(ilevel, which, None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1), dtype=np.int32)
        call("rh_get_level_table", ilevel, which, ptr(out, _lib._ip), n.value, ctypes.byref(n))
        return out[:n.value]

    # ---- level kernels ----
    def set_unew(self, head, num):
        call("rh_set_unew", head, num)

    def set_uold(self, head, num):
        call("rh_set_uold", head, num)

    def cmpdt(self, head, num, dx, p: HydroParams, gravity=(0.0, 0.0, 0.0)):
        g = np.ascontiguousarray(gravity, dtype=np.float64)
        out = [ctypes.c_double() for _ in range(5)]
        call("rh_cmpdt", head, num, dx, p.gamma, p.smallr, p.smallc ** 2, p.courant_factor, ptr(g),
             *[ctypes.byref(o) for o in out])
        return dict(zip(("mass", "ekin", "eint", "emag", "dt"), [o.value for o in out]))

    def godunov(self, head, num, ilevel, levelmin, levelmax, dt, dx, p: HydroParams, gravity=(0.0, 0.0, 0.0),
                precision=0):
        g = np.ascontiguousarray(gravity, dtype=np.float64)
        call("rh_set_precision", int(precision))
        call("rh_set_difmag", float(p.difmag))
        smax_cap = p.courant_factor * dx / dt
        call("rh_godunov", head, num, self.ngridmax, ilevel, levelmin, levelmax, p.gamma, p.smallr,
             p.smallc ** 2, dt, dx, smax_cap, p.slope_type, p.riemann_id, ptr(g))

    def sync(self):
        call("rh_device_sync")

    @staticmethod
    def set_brick_shape(bx, by, bz):
        call("rh_set_brick_shape", bx, by, bz)

    @staticmethod
    def brick_shape():
        v = [ctypes.c_int32() for _ in range(3)]
        call("rh_get_brick_shape", *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)


def newdt(dx, p: HydroParams, courant_dt, t=None, tout=None, nlanding=0, dtold=0.0):
    """dt rules of m_newdt_fine (pm/newdt_fine.f90:64-180) for one level without
    gravity/particles: dtnew = min(dx/smallc, courant); then, within nlanding
    steps of tout, land on it with dt = trem/nrem."""
    dtnew = dx / p.smallc
    dtnew = min(dtnew, courant_dt)
    if nlanding > 0 and tout is not None and t is not None:
        trem = tout - t
        if trem > 0:
            dtest = dtnew
            if dtold > 0:
                dtest = min(dtest, dtold)
            if trem < nlanding * dtest:
                nrem = max(1, int(np.ceil(trem / dtest - 1e-6)))
                dtnew = trem / float(nrem)
                if nrem == 1:
                    dtnew = trem + 1e-10 * max(trem, abs(tout))
    return dtnew


class UniformLevel:
    """A single uniform level (levelmin = levelmax) on the device."""

    def __init__(self, mesh: UniformMesh, cons_cells, params: HydroParams, periodic=True, bound_type=2,
                 ncachemax=0):
        self.mesh = mesh
        self.p = params
        nd = mesh.ndim
        self.hh = HipHydro(nd, mesh.noct, ncachemax)
        ckey_max, key_off, box_min, box_max = mesh.level_params()
        per = [1, 1, 1] if periodic else [0, 0, 0]
        if nd == 1:
            per = [int(periodic), 1, 1]
        self.hh.upload_level_params(ckey_max, key_off, box_min, box_max, per)
        self.hh.set_boundaries([bound_type] * 6)
        u = mesh.from_cells(cons_cells)
        self.hh.set_grid_device(u, u, mesh.grid_records())
        self.head, self.num, self.ilevel = 1, mesh.noct, mesh.level
        self.hh.insert_hash_all(1, mesh.noct)
        self.hh.build_nbor(1, mesh.noct, mesh.level)
        self.t = 0.0
        self.dtold = 0.0
        self.nstep = 0

    def courant(self):
        return self.hh.cmpdt(self.head, self.num, self.mesh.dx, self.p)

    def step(self, dt=None, tout=None, nlanding=0, precision=0):
        """One time step in the reference's order; returns dt used."""
        c = self.courant()
        if dt is None:
            dt = newdt(self.mesh.dx, self.p, c["dt"], self.t, tout, nlanding, self.dtold)
        self.hh.set_unew(self.head, self.num)
        self.hh.godunov(self.head, self.num, self.ilevel, self.ilevel, self.ilevel, dt, self.mesh.dx, self.p,
                        precision=precision)
        self.hh.set_uold(self.head, self.num)
        self.t += dt
        self.dtold = dt
        self.nstep += 1
        return dt

    def run_device(self, nsteps, dt_fixed=0.0, precision=0, use_graph=False, gravity=(0.0, 0.0, 0.0),
                   time_kernel=False, fused_courant=-1):
        """nsteps device-resident steps (dt reduced on the device).  Returns
        (dt_last, t_total, sums, godunov_ms) where godunov_ms is the mean
        Godunov launch duration (HIP events) when time_kernel is set.
        fused_courant: -1 automatic, 0 separate Courant kernel, 1 Courant
        reduction in the Godunov kernel's epilogue."""
        g = np.ascontiguousarray(gravity, dtype=np.float64)
        call("rh_set_precision", int(precision))
        call("rh_set_fused_courant", int(fused_courant))
        call("rh_set_difmag", float(self.p.difmag))
        dt_last = ctypes.c_double()
        t_tot = ctypes.c_double()
        sums = np.zeros(4)
        kms = ctypes.c_float(-1.0)
        call("rh_level_run", self.head, self.num, self.ilevel, self.mesh.dx, self.p.gamma, self.p.smallr,
             self.p.smallc ** 2, self.p.courant_factor, ptr(g), self.p.slope_type, self.p.riemann_id,
             float(dt_fixed), int(nsteps), int(use_graph), ctypes.byref(dt_last), ctypes.byref(t_tot), ptr(sums),
             ctypes.byref(kms) if time_kernel else None)
        self.t += t_tot.value
        self.nstep += nsteps
        return dt_last.value, t_tot.value, sums, (kms.value if time_kernel else None)

    def time_godunov(self, dt, nlaunch, precision=0):
        g = np.zeros(3)
        call("rh_set_precision", int(precision))
        ms = ctypes.c_float()
        call("rh_time_godunov", self.head, self.num, self.ilevel, self.mesh.dx, self.p.gamma, self.p.smallr,
             self.p.smallc ** 2, self.p.courant_factor, ptr(g), self.p.slope_type, self.p.riemann_id, float(dt),
             int(nlaunch), ctypes.byref(ms))
        return ms.value

    def state(self):
        return self.hh.transfer_grid_host(self.mesh.noct)

    def cells(self):
        return self.mesh.to_cells(self.state())


class AmrHierarchy:
    """A multi-level mesh (AmrMesh) on the device, stepped in the order of
    amr_step with nsubcycle=1 on every level (amr/amr_step.f90:276-588):
    newdt/courant and set_unew going down the levels, then, finest first,
    godunov_fine (refluxing into the coarser unew), set_uold and upload_fine.
    All levels share dt = min over levels (amr_step.f90:588)."""

    def __init__(self, amesh, uold, params: HydroParams, ncachemax=None, interpol_var=0, interpol_type=1,
                 precision=0, bricks=None, periodic=(1, 1, 1), bound_type=(2, 2, 2, 2, 2, 2), deterministic=False):
        """bricks: {level: (bx, by, bz)} brick shape per level (default: the
        library's current shape for every level).  periodic: per dimension;
        bound_type: (x low, x high, y low, ...) of the non-periodic faces, 1
        reflexive, 2 zero gradient.  deterministic: coarse-fine refluxes
        folded in a fixed order (rh_set_deterministic) instead of atomics."""
        self.am = amesh
        self.p = params
        self.precision = precision
        nc = ncachemax if ncachemax is not None else max(4 * amesh.ngrid, 10000)
        self.hh = HipHydro(3, amesh.ngrid, nc)
        ckey_max, key_off, box_min, box_max = amesh_level_params(amesh)
        self.hh.upload_level_params(ckey_max, key_off, box_min, box_max, list(periodic))
        self.hh.set_boundaries(list(bound_type))
        call("rh_set_deterministic", int(bool(deterministic)))
        call("rh_set_interpolation", int(interpol_var), int(interpol_type))
        self.interpol_var = interpol_var
        self.hh.set_grid_device(uold, uold, amesh.grid_records())
        self.hh.insert_hash_all(1, amesh.ngrid)
        for l in amesh.levels:
            if bricks is not None:
                HipHydro.set_brick_shape(*bricks[l])
            self.hh.build_nbor(amesh.head[l], amesh.noct[l], l)
        self.t = 0.0
        self.nstep = 0

    def courant(self, l):
        return self.hh.cmpdt(self.am.head[l], self.am.noct[l], self.am.dx(l), self.p)

    def step(self, dt=None):
        am = self.am
        dts = []
        for l in am.levels:
            c = self.courant(l)
            dts.append(min(am.dx(l) / self.p.smallc, c["dt"]))
            self.hh.set_unew(am.head[l], am.noct[l])
        if dt is None:
            dt = min(dts)
        for l in reversed(am.levels):
            self.hh.godunov(am.head[l], am.noct[l], l, am.levelmin, am.levelmax, dt, am.dx(l), self.p,
                            precision=self.precision)
            self.hh.set_uold(am.head[l], am.noct[l])
            if l < am.levelmax:
                call("rh_upload", am.head[l + 1], am.noct[l + 1], int(self.interpol_var == 1), self.p.gamma,
                     self.p.smallr, self.p.smallc ** 2)
        self.t += dt
        self.nstep += 1
        return dt

    def run_device(self, nsteps, precision=None, use_graph=False, gravity=(0.0, 0.0, 0.0)):
        """nsteps batched steps of all levels (rh_amr_prepare + rh_amr_run): one
        work-list Godunov launch over every level's bricks per step, dt kept
        on the device.  Returns (dt_next, t_advanced, sums, ms_per_step)."""
        am = self.am
        prec = self.precision if precision is None else precision
        head, noct = am.heads()
        g = np.ascontiguousarray(gravity, dtype=np.float64)
        call("rh_amr_prepare", am.levelmin, am.levelmax, ptr(head, _lib._ip), ptr(noct, _lib._ip), am.boxlen,
             self.p.gamma, self.p.smallr, self.p.smallc ** 2, self.p.courant_factor, ptr(g), self.p.slope_type,
             self.p.riemann_id)
        dt = ctypes.c_double()
        t = ctypes.c_double()
        sums = np.zeros(4)
        ms = ctypes.c_float()
        call("rh_amr_run", int(nsteps), int(prec), int(use_graph), ctypes.byref(dt), ctypes.byref(t), ptr(sums),
             ctypes.byref(ms))
        self.t += t.value
        self.nstep += nsteps
        return dt.value, t.value, sums, ms.value

    def state(self):
        return self.hh.transfer_grid_host(self.am.ngrid)


def amesh_level_params(amesh):
    return amesh.level_params()
