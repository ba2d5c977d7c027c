"""ctypes access to the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (the C restatement of the reference CPU path)
and, when it has been built in this tree, ``oracle/_ref/libramses_ref{1,3}d.so``
(the reference's own Fortran kernels).  Only tests/, ``__graft_entry__.smoke``
and the cpu_baseline leg of bench.py import this module, and only to check or
time; the product package ramses2_amd never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NVAR = 5

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class Params(ctypes.Structure):
    """ora_params_t (hydro_params_t subset, amr/amr_commons.f90:690-702)."""

    _fields_ = [("gamma", ctypes.c_double), ("smallr", ctypes.c_double),
                ("smallc", ctypes.c_double), ("courant_factor", ctypes.c_double),
                ("smax_cap", ctypes.c_double), ("difmag", ctypes.c_double),
                ("slope_type", ctypes.c_int), ("riemann", ctypes.c_int)]


def make_params(gamma=1.4, smallr=1e-10, smallc=1e-10, courant_factor=0.5,
                slope_type=1, riemann=3, difmag=0.0, smax_cap=1.7976931348623157e308):
    return Params(gamma, smallr, smallc, courant_factor, smax_cap, difmag,
                  slope_type, riemann)


class Mesh(ctypes.Structure):
    """ora_mesh_t; the numpy arrays it points to are kept alive on the object."""

    _fields_ = [("ndim", ctypes.c_int), ("nvar", ctypes.c_int), ("ngrid", ctypes.c_int),
                ("ckey", _ip), ("lev", _ip), ("refined", _u8p),
                ("uold", _dp), ("unew", _dp),
                ("levelmin", ctypes.c_int), ("nlevelmax", ctypes.c_int),
                ("head", _ip), ("noct", _ip), ("boxlen", ctypes.c_double),
                ("periodic", ctypes.c_int * 3), ("bound_type", (ctypes.c_int * 2) * 3),
                ("interpol_var", ctypes.c_int), ("interpol_type", ctypes.c_int),
                ("constant_gravity", ctypes.c_double * 3),
                ("box_lev", ctypes.c_int), ("box_ckey_min", ctypes.c_int * 3),
                ("box_ckey_max", ctypes.c_int * 3)]


def _ptr(a, t):
    return a.ctypes.data_as(t)


def build():
    """Compile liboracle.so (gcc); the reference build needs /root/reference."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def build_ref():
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        for name in ("ora_riemann_llf", "ora_riemann_hll", "ora_riemann_hllc"):
            getattr(L, name).argtypes = [_dp, _dp, _dp, ctypes.POINTER(Params)]
        L.ora_unsplit.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                  ctypes.c_double, ctypes.c_double, ctypes.POINTER(Params)]
        L.ora_cmpdt.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double,
                                ctypes.POINTER(Params), _dp]
        L.ora_interpol_hydro.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double]
        L.ora_godunov_fine.argtypes = [ctypes.POINTER(Mesh), ctypes.c_int, ctypes.c_double,
                                       ctypes.POINTER(Params)]
        L.ora_godunov_fine.restype = ctypes.c_int
        L.ora_courant_fine.argtypes = [ctypes.POINTER(Mesh), ctypes.c_int,
                                       ctypes.POINTER(Params), ctypes.c_int, _dp]
        L.ora_upload_fine.argtypes = [ctypes.POINTER(Mesh), ctypes.c_int, ctypes.c_double]
        L.ora_upload_fine.restype = ctypes.c_int
        L.ora_set_unew.argtypes = [ctypes.POINTER(Mesh), ctypes.c_int]
        L.ora_set_uold.argtypes = [ctypes.POINTER(Mesh), ctypes.c_int]
        L.ora_godunov_uniform_omp.argtypes = [ctypes.POINTER(Mesh), ctypes.c_int,
                                              ctypes.c_double, ctypes.POINTER(Params),
                                              _ip, ctypes.c_int]
        L.ora_godunov_uniform_omp.restype = ctypes.c_int
        L.ora_probe_get_grid.argtypes = [ctypes.POINTER(Mesh), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int64), _dp, _u8p]
        L.ora_probe_get_grid.restype = ctypes.c_int
        _lib = L
    return _lib


def ref_lib(ndim):
    """The flang-built reference kernels for ``ndim`` (None if not built)."""
    path = os.path.join(HERE, "_ref", "libramses_ref%dd.so" % ndim)
    if not os.path.exists(path):
        return None
    # Lazy binding: domain_hilbert.f90 (pulled in by amr_commons) references
    # wallclock() from amr/timer.f90, which the hydro kernels never call.
    L = ctypes.CDLL(path, mode=os.RTLD_LAZY)
    d = ctypes.c_double
    L.ref_unsplit.argtypes = [ctypes.c_int, _dp, _dp, _dp, d, d, d, d, d, d, d,
                              ctypes.c_int, ctypes.c_int]
    L.ref_riemann.argtypes = [ctypes.c_int, _dp, _dp, _dp, d, d, d, d]
    L.ref_cmpdt.argtypes = [_dp, _dp, d, d, d, d, d, _dp]
    L.ref_interpol_hydro.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, d]
    L.ref_ndim.restype = ctypes.c_int
    return L


# --------------------------------------------------------------------------
# Thin numpy wrappers
# --------------------------------------------------------------------------
def stencil_shape(ndim, nn=2):
    n = nn + 4
    return (n, n if ndim > 1 else 1, n if ndim > 2 else 1)


def flux_shape(ndim, nn=2):
    n = nn + 1
    return (n, n if ndim > 1 else 1, n if ndim > 2 else 1)


def unsplit(ndim, uin, gin, dx, dt, p, nn=2):
    """uin: Fortran-ordered (ni,nj,nk,nvar); returns flux (fi,fj,fk,nvar,ndim)."""
    uin = np.asfortranarray(uin, dtype=np.float64)
    gin = np.asfortranarray(gin, dtype=np.float64)
    flux = np.zeros(flux_shape(ndim, nn) + (NVAR, ndim), dtype=np.float64, order="F")
    pp = Params.from_buffer_copy(p)
    lib().ora_unsplit(ndim, nn, _ptr(uin, _dp), _ptr(gin, _dp), _ptr(flux, _dp), dx, dt,
                      ctypes.byref(pp))
    return flux


def riemann(solver, ql, qr, p):
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    f = np.zeros(6)
    fn = {1: lib().ora_riemann_llf, 2: lib().ora_riemann_hll, 3: lib().ora_riemann_hllc}[solver]
    fn(_ptr(ql, _dp), _ptr(qr, _dp), _ptr(f, _dp), ctypes.byref(p))
    return f


def cmpdt(ndim, uu, gg, dx, p):
    uu = np.array(uu, dtype=np.float64)
    gg = np.zeros(3) if gg is None else np.array(gg, dtype=np.float64)
    dt = ctypes.c_double()
    lib().ora_cmpdt(ndim, _ptr(uu, _dp), _ptr(gg, _dp), dx, ctypes.byref(p), ctypes.byref(dt))
    return dt.value, uu


def interpol_hydro(ndim, u1, interpol_var, interpol_type, smallr=1e-10):
    u1 = np.array(u1, dtype=np.float64, order="F")
    u2 = np.zeros((2 ** ndim, NVAR), dtype=np.float64, order="F")
    lib().ora_interpol_hydro(ndim, _ptr(u1, _dp), _ptr(u2, _dp), interpol_var, interpol_type,
                             smallr)
    return u2, u1


class OracleMesh:
    """Owns numpy arrays of one mesh and exposes them as an ora_mesh_t."""

    def __init__(self, ndim, ckey, lev, refined, uold, head, noct, levelmin, boxlen=1.0,
                 periodic=(1, 1, 1), bound_type=((2, 2), (2, 2), (2, 2)),
                 interpol_var=0, interpol_type=1, gravity=(0.0, 0.0, 0.0), unew=None, box=None):
        """box = (lev, ckey_min[3], ckey_max[3]) for a computational box smaller
        than the full 2^(lev-1)-oct level (box_ckey_min/max of the reference)."""
        ngrid = len(lev)
        self.ndim = ndim
        self.ckey = np.zeros((ngrid, 3), dtype=np.int32)
        self.ckey[:, :ckey.shape[1]] = ckey
        self.lev = np.ascontiguousarray(lev, dtype=np.int32)
        self.refined = np.zeros((ngrid, 8), dtype=np.uint8)
        self.refined[:, :refined.shape[1]] = refined
        self.uold = np.ascontiguousarray(uold, dtype=np.float64)
        self.unew = np.array(uold if unew is None else unew, dtype=np.float64)
        nl = len(head)
        self.head = np.ascontiguousarray(head, dtype=np.int32)
        self.noct = np.ascontiguousarray(noct, dtype=np.int32)
        m = Mesh()
        m.ndim, m.nvar, m.ngrid = ndim, NVAR, ngrid
        m.ckey = _ptr(self.ckey, _ip)
        m.lev = _ptr(self.lev, _ip)
        m.refined = _ptr(self.refined, _u8p)
        m.uold = _ptr(self.uold, _dp)
        m.unew = _ptr(self.unew, _dp)
        m.levelmin, m.nlevelmax = levelmin, nl
        m.head = _ptr(self.head, _ip)
        m.noct = _ptr(self.noct, _ip)
        m.boxlen = boxlen
        for d in range(3):
            m.periodic[d] = int(periodic[d])
            m.bound_type[d][0] = int(bound_type[d][0])
            m.bound_type[d][1] = int(bound_type[d][1])
            m.constant_gravity[d] = gravity[d]
        m.interpol_var, m.interpol_type = interpol_var, interpol_type
        if box is not None:
            m.box_lev = int(box[0])
            for d in range(3):
                m.box_ckey_min[d] = int(box[1][d])
                m.box_ckey_max[d] = int(box[2][d])
        self.m = m

    def godunov_fine(self, ilevel, dt, p):
        rc = lib().ora_godunov_fine(ctypes.byref(self.m), ilevel, dt, ctypes.byref(p))
        if rc != 0:
            raise RuntimeError("ora_godunov_fine failed with code %d" % rc)

    def courant_fine(self, ilevel, p, nthreads=1):
        out = np.zeros(5)
        lib().ora_courant_fine(ctypes.byref(self.m), ilevel, ctypes.byref(p), nthreads,
                               _ptr(out, _dp))
        return out

    def set_unew(self, ilevel):
        lib().ora_set_unew(ctypes.byref(self.m), ilevel)

    def upload_fine(self, ilevel, smallr=1e-10):
        rc = lib().ora_upload_fine(ctypes.byref(self.m), ilevel, smallr)
        if rc != 0:
            raise RuntimeError("ora_upload_fine: parent cell missing")

    def set_uold(self, ilevel):
        lib().ora_set_uold(ctypes.byref(self.m), ilevel)

    def probe_get_grid(self, lev, key):
        """get_grid of one key: (code, u[nvar, ttd], refined[ttd]); code > 0 an
        oct of the mesh, -1 a boundary ghost, 0 missing."""
        ttd = 2 ** self.ndim
        k = (ctypes.c_int64 * 3)(*([int(x) for x in key] + [0] * (3 - len(key))))
        u = np.zeros((NVAR, ttd))
        r = np.zeros(ttd, dtype=np.uint8)
        c = lib().ora_probe_get_grid(ctypes.byref(self.m), lev, k, _ptr(u, _dp), _ptr(r, _u8p))
        return c, u, r

    def hydro_flag(self, ilevel, gamma, smallr, err_grad_d, err_grad_p, floor_d=1e-10, floor_p=1e-10, flag1=None):
        """ora_hydro_flag: flag1 (ngrid, 2^ndim) int32, OR-ed into the given
        array (zeros when None); returns (flag1, newly flagged count)."""
        ttd = 2 ** self.ndim
        f = np.zeros((self.m.ngrid, ttd), dtype=np.int32) if flag1 is None else np.ascontiguousarray(flag1, np.int32)
        L = lib()
        L.ora_hydro_flag.restype = ctypes.c_int
        n = L.ora_hydro_flag(ctypes.byref(self.m), ilevel, ctypes.c_double(gamma), ctypes.c_double(smallr),
                             ctypes.c_double(err_grad_d), ctypes.c_double(err_grad_p), ctypes.c_double(floor_d),
                             ctypes.c_double(floor_p), _ptr(f, _ip))
        if n < 0:
            raise RuntimeError("ora_hydro_flag: a neighbour's parent cell is missing")
        return f, n

    def godunov_uniform_omp(self, ilevel, dt, p, octmap, nthreads):
        octmap = np.ascontiguousarray(octmap, dtype=np.int32)
        rc = lib().ora_godunov_uniform_omp(ctypes.byref(self.m), ilevel, dt, ctypes.byref(p),
                                           _ptr(octmap, _ip), nthreads)
        if rc != 0:
            raise RuntimeError("ora_godunov_uniform_omp failed with code %d" % rc)


# --------------------------------------------------------------------------
# The reference's own level routines (oracle/_ref/libramses_reflev{1,3}d.so)
# --------------------------------------------------------------------------
_reflev = {}


def reflev_lib(ndim):
    """The flang-built reference level routines for ``ndim`` (None if not built)."""
    if ndim in _reflev:
        return _reflev[ndim]
    path = os.path.join(HERE, "_ref", "libramses_reflev%dd.so" % ndim)
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path, mode=os.RTLD_LAZY)
    d, i = ctypes.c_double, ctypes.c_int
    i3 = ctypes.c_int * 3
    L.reflev_ndim.restype = i
    L.reflev_setup.argtypes = [i, _ip, _ip, ctypes.POINTER(ctypes.c_int8), _dp, _dp, i, i, _ip, _ip, d,
                               i3, ctypes.c_int * 6, i, i3, i3, i, i, ctypes.c_double * 3,
                               d, d, d, d, i, i, d]
    L.reflev_godunov.argtypes = [i, d]
    L.reflev_courant.argtypes = [i, _dp]
    L.reflev_upload.argtypes = [i]
    L.reflev_set_unew.argtypes = [i]
    L.reflev_set_uold.argtypes = [i]
    L.reflev_get.argtypes = [i, _dp, _dp]
    L.reflev_get_grid.argtypes = [i, i3, _dp, ctypes.POINTER(ctypes.c_int8)]
    L.reflev_get_grid.restype = i
    _reflev[ndim] = L
    return L


class RefLevel:
    """The mesh of an OracleMesh loaded into the reference's own types; its
    methods run the reference's godunov_fine / courant_fine / upload_fine /
    set_unew / set_uold (one instance at a time per ndim: the library keeps
    the mesh in module variables).

    godunov_fine and upload_fine close the reference's software cache on
    exit, which destages every boundary ghost oct through destage()
    (amr/task_manager.f90:974-1069).  That file's module closure takes in the
    whole program, including lightcone/lightcone.f90, which flang rejects, so
    it is not in the library: both are refused on meshes with a non-periodic
    dimension, and after probe_get_grid (whose ghosts stay in the cache)."""

    def __init__(self, om, p):
        L = reflev_lib(om.ndim)
        if L is None:
            raise RuntimeError("reference level library not built (make -C oracle reflev)")
        self.L, self.ngrid = L, om.m.ngrid
        m = om.m
        self.ndim = om.ndim
        self.periodic = all(m.periodic[d] for d in range(om.ndim))
        self.probed = False
        self._ckey = np.ascontiguousarray(om.ckey, dtype=np.int32)
        self._ref = np.ascontiguousarray(om.refined.astype(np.int8))
        self._uold = np.ascontiguousarray(om.uold, dtype=np.float64)
        self._unew = np.ascontiguousarray(om.unew, dtype=np.float64)
        i3 = ctypes.c_int * 3
        per = i3(*[m.periodic[k] for k in range(3)])
        bt = (ctypes.c_int * 6)(*[m.bound_type[k][s] for k in range(3) for s in range(2)])
        bmin = i3(*[m.box_ckey_min[k] for k in range(3)])
        bmax = i3(*[m.box_ckey_max[k] for k in range(3)])
        g = (ctypes.c_double * 3)(*[m.constant_gravity[k] for k in range(3)])
        L.reflev_setup(m.ngrid, _ptr(self._ckey, _ip), _ptr(om.lev, _ip),
                       self._ref.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                       _ptr(self._uold, _dp), _ptr(self._unew, _dp), m.levelmin, m.nlevelmax,
                       _ptr(om.head, _ip), _ptr(om.noct, _ip), m.boxlen, per, bt, m.box_lev, bmin, bmax,
                       m.interpol_var, m.interpol_type, g, p.gamma, p.smallr, p.smallc, p.courant_factor,
                       p.slope_type, p.riemann, p.difmag)

    def _cache_ok(self):
        if not self.periodic or self.probed:
            raise RuntimeError("closing the reference's cache needs destage (amr/task_manager.f90)")

    def godunov_fine(self, ilevel, dt):
        self._cache_ok()
        self.L.reflev_godunov(ilevel, dt)

    def courant_fine(self, ilevel):
        out = np.zeros(5)
        self.L.reflev_courant(ilevel, _ptr(out, _dp))
        return out

    def upload_fine(self, ilevel):
        self._cache_ok()
        self.L.reflev_upload(ilevel)

    def probe_get_grid(self, lev, key):
        """The reference's get_grid of one key with godunov_fine's boundary
        initialiser: (child, u[nvar, ttd], refined[ttd]); child > ngrid is a
        boundary ghost oct built in the cache, 0 missing."""
        ttd = 2 ** self.ndim
        k = (ctypes.c_int * 3)(*([int(x) for x in key] + [0] * (3 - len(key))))
        u = np.zeros((NVAR, ttd))
        r = np.zeros(ttd, dtype=np.int8)
        self.probed = True
        c = self.L.reflev_get_grid(lev, k, _ptr(u, _dp), r.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
        return c, u, r.astype(np.uint8)

    def hydro_flag(self, ilevel, err_grad_d, err_grad_p, floor_d=1e-10, floor_p=1e-10):
        """The reference's hydro_flag on level ilevel (flag1 cleared first):
        flag1 (ngrid, 2^ndim) int32."""
        self._cache_ok()
        ttd = 2 ** self.ndim
        f = np.zeros((self.ngrid, ttd), dtype=np.int32)
        d = ctypes.c_double
        self.L.reflev_hydro_flag(ilevel, d(err_grad_d), d(err_grad_p), d(floor_d), d(floor_p), self.ngrid,
                                 _ptr(f, _ip))
        return f

    def set_unew(self, ilevel):
        self.L.reflev_set_unew(ilevel)

    def set_uold(self, ilevel):
        self.L.reflev_set_uold(ilevel)

    def state(self):
        """(uold, unew) of the real octs."""
        uo = np.zeros((self.ngrid, NVAR, self._uold.shape[2]))
        un = np.zeros_like(uo)
        self.L.reflev_get(self.ngrid, _ptr(uo, _dp), _ptr(un, _dp))
        return uo, un
