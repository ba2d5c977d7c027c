"""ctypes signatures of the extension entry points (include/ramses2_hip_ext.h)."""
import ctypes

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_fp = ctypes.POINTER(ctypes.c_float)
_c = ctypes.c_int
_d = ctypes.c_double

SIGNATURES = {
    "rh_set_precision": [_c],
    "rh_set_fused_courant": [_c],
    "rh_set_brick_shape": [_c, _c, _c],
    "rh_set_deterministic": [_c],
    "rh_set_specialized": [_c],
    "rh_set_difmag": [_d],
    "rh_get_brick_shape": [_ip, _ip, _ip],
    "rh_timer_start": [],
    "rh_timer_stop": [_fp],
    "rh_level_run": [_c, _c, _c, _d, _d, _d, _d, _d, _dp, _c, _c, _d, _c, _c, _dp, _dp, _dp, _fp],
    "rh_copy_probe": [ctypes.c_int64, _fp],
    "rh_time_godunov": [_c, _c, _c, _d, _d, _d, _d, _d, _dp, _c, _c, _d, _c, _fp],
    "rh_amr_prepare": [_c, _c, _ip, _ip, _d, _d, _d, _d, _d, _dp, _c, _c],
    "rh_amr_run": [_c, _c, _c, _dp, _dp, _dp, _fp],
    "rh_ctx_create": [_ip],
    "rh_ctx_select": [_c],
    "rh_ctx_current": [_ip],
    "rh_ctx_destroy": [_c],
    "rh_ctx_footprint": [_c, ctypes.POINTER(ctypes.c_int64)],
    "rh_ctx_clone_mesh": [_c],
    "rh_cache_release": [ctypes.POINTER(ctypes.c_int64)],
    "rh_estimate_uniform_bytes": [_c, _c, _c, ctypes.POINTER(ctypes.c_int64)],
    "rh_init_ic": [_c, _c, _c, _c, _dp, _d],
    "rh_job_prepare": [_c, _c, _c, _d, _d, _d, _d, _d, _dp, _c, _c],
    "rh_batch_step": [_ip, _c, _c, _c],
    "rh_job_status": [_c, _dp],
    "rh_fence_record": [_ip, _c, _ip],
    "rh_fence_wait": [_c, _dp],
    "rh_get_level_table": [_c, _c, _ip, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)],
    "rh_brick_lattice": [_ip, _c, _ip, _ip, _c, _c, _c, _ip, _dp],
}


def bind(L):
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
