"""ctypes signatures of the extension entry points (include/ramses2_hip_ext.h)."""
import ctypes

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_fp = ctypes.POINTER(ctypes.c_float)
_c = ctypes.c_int
_d = ctypes.c_double

SIGNATURES = {
    "rh_set_precision": [_c],
    "rh_set_brick_shape": [_c, _c, _c],
    "rh_get_brick_shape": [_ip, _ip, _ip],
    "rh_timer_start": [],
    "rh_timer_stop": [_fp],
    "rh_level_run": [_c, _c, _c, _d, _d, _d, _d, _d, _dp, _c, _c, _d, _c, _c, _dp, _dp, _dp, _fp],
    "rh_copy_probe": [ctypes.c_int64, _fp],
    "rh_time_godunov": [_c, _c, _c, _d, _d, _d, _d, _d, _dp, _c, _c, _d, _c, _fp],
}


def bind(L):
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
