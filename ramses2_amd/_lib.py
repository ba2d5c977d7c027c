"""ctypes binding of the C-ABI declared in include/ramses2_hip.h.

The shared library is loaded from ramses2_amd/lib/ (built in-tree by
ramses2_amd/build.py).  There is no fallback: if the library is missing or a
call fails, an exception is raised.
"""
import ctypes
import os

from . import build as _build

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_c = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p

# name -> argtypes (restype int unless stated)
SIGNATURES = {
    "rh_init": [],
    "rh_finalize": [],
    "rh_alloc_amr": [_c, _c, _c, _c, _c],
    "rh_upload_level_params": [_ip, _lp, _ip, _ip, _ip, _c],
    "rh_set_boundaries": [_ip],
    "rh_set_grid_device": [_dp, _dp, _vp, _vp, _c, _c, _c],
    "rh_transfer_grid_host": [_dp, _vp, _c, _c, _c],
    "rh_transfer_unew_host": [_dp, _c, _c, _c],
    "rh_insert_hash_all": [_c, _c, _c],
    "rh_build_nbor": [_c, _c, _c, _c, _i64, _ip, _ip, _ip],
    "rh_set_unew": [_c, _c],
    "rh_set_uold": [_c, _c],
    "rh_cmpdt": [_c, _c, _d, _d, _d, _d, _d, _dp, _dp, _dp, _dp, _dp, _dp],
    "rh_cmpdt_noalloc": [_c, _c, _d, _d, _d, _d, _d, _dp, _dp, _dp, _dp, _dp, _dp],
    "rh_godunov": [_c, _c, _c, _c, _c, _c, _d, _d, _d, _d, _d, _d, _c, _c, _dp],
    "rh_device_sync": [],
    "rh_upload": [_c, _c, _c, _d, _d, _d],
    "rh_set_interpolation": [_c, _c],
    "rh_upload_flag1": [_ip, _c],
    "rh_transfer_flag1_host": [_ip, _c],
    "rh_reset_flag1": [_c, _c],
    "rh_count_flag1": [_c, _c, _ip],
    "rh_hydro_flag": [_c, _c, _d, _d, _d, _d, _d, _d, _d],
    "rh_last_error": None,
}

_lib = None


class HipError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("RAMSES2_HIP_LIB") or _build.lib_path()
    if not os.path.exists(path):
        raise ImportError("ramses2_amd: %s is not built; run ramses2_amd/build.py "
                          "(or __graft_entry__.build())" % path)
    L = ctypes.CDLL(path)
    L.rh_last_error.restype = ctypes.c_char_p
    L.rh_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        if args is None:
            continue
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    from . import _ext_sigs
    _ext_sigs.bind(L)
    _lib = L
    return L


def call(name, *args):
    """Call an rh_* entry point and raise HipError on a nonzero return code."""
    L = load()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.rh_last_error().decode(errors="replace")
        raise HipError("%s failed (code %d): %s" % (name, rc, msg))
    return rc


def ptr(a, t=_dp):
    return a.ctypes.data_as(t)
