"""ctypes binding of libpfast.so (the C ABI declared in include/pfast.h).

The library is loaded from this package directory.  There is no fallback: if it
is missing or fails to load, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfast.so")

PFAST_OK = 0
PFAST_ERR_ARG = -1
PFAST_ERR_HIP = -2
PFAST_ERR_NOMEM = -3
PFAST_ERR_FATAL = -4
PFAST_FWD_ONLY = 1
PFAST_FWD_ADJ = 2
PFAST_REQ_TABLE = 1
PFAST_REQ_GRADIENT = 2
REQ_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class PfastError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpfast error {code}: {msg}")
        self.code = code


class FatalError(PfastError):
    """A condition on which the reference program calls exit(1)."""


dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
d, i, vpp = C.c_double, C.c_int, C.c_void_p
pint, pdbl, plong = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_long)

# name -> argtypes (all return int status)
_SIGS = {
    "pfast_lbfgs_create": [vpp, C.c_long, i, C.POINTER(vpp)],
    "pfast_lbfgs_destroy": [vpp],
    "pfast_lbfgs_direction": [vpp, i, dp, dp, dp, d, pdbl],
    "pfast_lbfgs_dot": [vpp, dp, dp, pdbl],
    "pfast_version": [],
    "pfast_device_count": [pint],
    "pfast_fsm3d": [dp, dp, ip, i, i, d, d, d, i, i, i, i, vpp],
    "pfast_fsm3d_noalloc": [dp, dp, ip, i, i, d, d, d, i, i, i, i, vpp],
    "pfast_adjoint_fsm3d": [dp, dp, vpp, vpp, vpp, ip, ip, i, i, d, d, d, i, i, i, i, pdbl, vpp],
    "pfast_rt_fsm2d": [dp, dp, vpp, vpp, vpp, ip, vpp, vpp, i, i, i, d, d, i, i, i, vpp, vpp],
    "pfast_tadjoint_fsm2d": [dp, dp, vpp, d, vpp, vpp, ip, ip, i, i, d, d, i, i, i, pdbl, vpp],
    "pfast_radjoint_fsm2d": [dp, vpp, dp, vpp, d, vpp, vpp, ip, ip, i, i, i, d, d, i, i, i, pdbl, vpp],
    "pfast_ctx_create": [i, C.POINTER(vpp)],
    "pfast_ctx_destroy": [vpp],
    "pfast_ctx_sync": [vpp],
    "pfast_ctx_timing": [vpp, i, pdbl, plong],
    "pfast_ctx_timing_enable": [vpp, i],
    "pfast_ctx_timing_reset": [vpp],
    "pfast_ctx_set_sweep": [vpp, i],
    "pfast_ctx_set_sweep2d": [vpp, i],
    "pfast_ctx_set_streams": [vpp, i],
    "pfast_ctx_set_tile_skip": [vpp, i],
    "pfast_ctx_set_paired": [vpp, i],
    "pfast_ctx_set_dataflow": [vpp, i],
    "pfast_ctx_set_check_interval": [vpp, i],
    "pfast_ctx_set_graphs": [vpp, i],
    "pfast_ctx_set_source_order": [vpp, i],
    "pfast_source_order": [i, vpp, C.c_double, C.c_double, C.c_double, i, vpp],
    "pfast_model3d_create": [vpp, dp, i, i, i, d, d, d, C.POINTER(vpp)],
    "pfast_model3d_set_vp": [vpp, dp],
    "pfast_model3d_destroy": [vpp],
    "pfast_shots3d_create": [vpp, i, ip, ip, vpp, vpp, vpp, C.POINTER(vpp)],
    "pfast_shots3d_destroy": [vpp],
    "pfast_shots3d_run": [vpp, i, i],
    "pfast_shots3d_set_source_init": [vpp, i],
    "pfast_shots3d_autotune": [vpp, vpp, i, pint, vpp],
    "pfast_shots3d_autotune_schedule": [vpp, pint, pint, vpp],
    "pfast_shots3d_gradient": [vpp, vpp, pdbl],
    "pfast_shots3d_picks": [vpp, dp],
    "pfast_shots3d_field": [vpp, i, i, dp],
    "pfast_shots3d_cycles": [vpp, ip],
    "pfast_shots3d_plain_source": [vpp, i, i, pdbl, vpp, vpp, vpp],
    "pfast_ctx_work": [vpp, i, pdbl],
    "pfast_ctx_work_swept": [vpp, i, pdbl],
    "pfast_model2d_create": [vpp, dp, i, i, d, d, vpp, i, C.POINTER(vpp)],
    "pfast_model2d_set_vp": [vpp, dp],
    "pfast_model2d_destroy": [vpp],
    "pfast_shots2d_create": [vpp, i, ip, ip, vpp, vpp, vpp, vpp, C.POINTER(vpp)],
    "pfast_shots2d_destroy": [vpp],
    "pfast_shots2d_run": [vpp, i, i],
    "pfast_shots2d_gradient": [vpp, vpp, vpp, vpp, vpp, vpp],
    "pfast_shots2d_field": [vpp, i, i, i, dp],
    "pfast_shots2d_info": [vpp, vpp, vpp],
    "pfast_regu3d": [dp, dp, d, d, d, d, d, d, i, i, i],
    "pfast_regu2d": [dp, dp, d, d, d, d, i, i],
    "pfast_server_create": [i, C.c_size_t, C.POINTER(vpp)],
    "pfast_server_destroy": [vpp],
    "pfast_server_add_model3d": [vpp, dp, i, i, i, d, d, d, pint],
    "pfast_server_add_model2d": [vpp, dp, i, i, d, d, vpp, i, pint],
    "pfast_server_remove_model": [vpp, i],
    "pfast_server_submit3d": [vpp, i, i, i, ip, ip, vpp, vpp, vpp, i, plong],
    "pfast_server_submit2d": [vpp, i, i, i, ip, ip, vpp, vpp, vpp, vpp, i, plong],
    "pfast_server_progress": [vpp, i, pint],
    "pfast_server_status": [vpp, C.c_long, pint, vpp],
    "pfast_server_cycles": [vpp, C.c_long, vpp],
    "pfast_server_gradient": [vpp, C.c_long, vpp, vpp, vpp],
    "pfast_server_field": [vpp, C.c_long, i, dp],
    "pfast_server_release": [vpp, C.c_long],
    "pfast_server_footprint": [vpp, C.c_long, C.POINTER(C.c_size_t)],
    "pfast_server_info": [vpp, C.POINTER(C.c_size_t), pint, pint, plong, pdbl],
}
# entry points returning double rather than a status code
_DSIGS = {
    "pfast_beta_hs": [dp, dp, dp, C.c_long],
    "pfast_beta_cgdescent": [dp, dp, dp, C.c_long],
    "pfast_gd2sk": [dp, dp, dp, d, dp, dp, dp, i, i, C.c_long],
    "pfast_zoom": [d, d, d, d, d, d],
    "pfast_cubic_interp": [d, d, d, d, d, d],
    "pfast_magnitude_of": [dp, C.c_long],
    "pfast_mean_of": [dp, C.c_long],
}

_lib = None


def lib():
    """The loaded libpfast (raises if the library is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PfastError(PFAST_ERR_HIP, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        for name, args in _DSIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_double
        L.pfast_last_error.argtypes = []
        L.pfast_last_error.restype = C.c_char_p
        L.pfast_sweep_name.argtypes = [i]
        L.pfast_sweep_name.restype = C.c_char_p
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS) + list(_DSIGS) + ["pfast_last_error", "pfast_sweep_name"]


def check(rc):
    if rc != PFAST_OK:
        msg = lib().pfast_last_error().decode(errors="replace")
        if rc == PFAST_ERR_FATAL:
            raise FatalError(rc, msg)
        raise PfastError(rc, msg)
    return rc


def ptr(a):
    """Address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(vpp)
