"""ctypes binding of libfd23.so (the C ABI in include/fd23.h and
include/fd23_serve.h).

The library is built in-tree by ``fd23_amd.build`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing:
there is no CPU fallback for the time-stepping path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfd23.so")

OK = 0
EINVAL, ENOMEM, EHIP, EIO, ESTAB = -1, -2, -3, -4, -5

WAVE_PSV, WAVE_SH = 0, 1
STORE_F32, STORE_F16, STORE_BF16, STORE_F64 = 0, 1, 2, 3
SRC_LINE, SRC_P, SRC_FORCE = 0, 1, 2
TILE_TILED, TILE_FUSED, TILE_BLOCK = 0, 1, 2

# exported symbols and their signatures (kept in step with include/fd23.h)
c_int, c_dbl, c_flt, c_vp, c_sz = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_void_p, ctypes.c_size_t
P_int = ctypes.POINTER(ctypes.c_int)
P_dbl = ctypes.POINTER(ctypes.c_double)
P_flt = ctypes.POINTER(ctypes.c_float)


class Desc(ctypes.Structure):
    _fields_ = [("wave", c_int), ("storage", c_int), ("nx", c_int), ("nz", c_int), ("npml", c_int),
                ("topabsorb", c_int), ("nshots", c_int), ("max_frames", c_int), ("max_rec", c_int),
                ("stream", c_vp)]


class ServerDesc(ctypes.Structure):
    _fields_ = [("hbm_budget", c_sz), ("max_active", c_int), ("chunk_steps", c_int), ("max_batch_shots", c_int),
                ("autotune", c_int), ("spare_shots", c_int)]


class Request(ctypes.Structure):
    _fields_ = [("wave", c_int), ("storage", c_int), ("nx", c_int), ("nz", c_int), ("npml", c_int),
                ("topabsorb", c_int), ("nshots", c_int), ("nt", c_int), ("itrecord", c_int),
                ("dx", c_dbl), ("dt", c_dbl), ("material", ctypes.POINTER(c_vp)), ("cpml", c_vp),
                ("src_kind", c_int), ("xs", c_vp), ("zs", c_vp), ("stf", c_vp), ("lsrc", c_int), ("M", c_vp),
                ("nrec", c_int), ("rx", c_vp), ("rz", c_vp), ("box", c_int), ("box_x0", c_int),
                ("box_z0", c_int), ("box_bx", c_int), ("box_bz", c_int), ("box_incident", c_int),
                ("box_angle", c_dbl), ("box_f0", c_dbl), ("box_t0", c_dbl), ("box_amp", c_dbl),
                ("box_vp", c_dbl), ("box_vs", c_dbl), ("box_rho", c_dbl), ("model_key", ctypes.c_ulonglong),
                ("model", ctypes.POINTER(c_vp))]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("failed", ctypes.c_longlong),
                ("queued", ctypes.c_longlong), ("running", ctypes.c_longlong), ("launches", ctypes.c_longlong),
                ("merged", ctypes.c_longlong), ("peak_active", ctypes.c_longlong), ("bytes_in_use", c_sz),
                ("budget", c_sz), ("cell_updates", c_dbl), ("joined", ctypes.c_longlong)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

SIGNATURES = {
    "fd23_ctx_create": (c_int, [ctypes.POINTER(Desc), ctypes.POINTER(c_vp)]),
    "fd23_ctx_destroy": (c_int, [c_vp]),
    "fd23_last_error": (ctypes.c_char_p, []),
    "fd23_ctx_device_bytes": (c_sz, [c_vp]),
    "fd23_desc_device_bytes": (c_sz, [ctypes.POINTER(Desc)]),
    "fd23_set_material": (c_int, [c_vp, ctypes.POINTER(c_vp)]),
    "fd23_set_model": (c_int, [c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, ctypes.POINTER(c_dbl)]),
    "fd23_set_model_noalloc": (c_int, [c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, ctypes.POINTER(c_dbl)]),
    "fd23_set_model_noalloc_capacity": (ctypes.c_size_t, []),
    "fd23_get_material": (c_int, [c_vp, c_int, c_vp]),
    "fd23_set_model_file": (c_int, [c_vp, ctypes.c_char_p, c_dbl, c_dbl, ctypes.POINTER(c_dbl)]),
    "fd23_set_cpml": (c_int, [c_vp, c_vp]),
    "fd23_set_source": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_int, c_vp]),
    "fd23_set_receivers": (c_int, [c_vp, c_int, c_int, c_vp, c_vp, c_int]),
    "fd23_set_shot_slot": (c_int, [c_vp, c_int, c_int, c_int]),
    "fd23_reset_shot": (c_int, [c_vp, c_int]),
    "fd23_step_stress": (c_int, [c_vp]),
    "fd23_inject": (c_int, [c_vp, c_int]),
    "fd23_step_velocity": (c_int, [c_vp]),
    "fd23_record": (c_int, [c_vp, c_int, c_int]),
    "fd23_run": (c_int, [c_vp, c_int, c_int, c_int]),
    "fd23_prepare": (c_int, [c_vp, c_int, c_int]),
    "fd23_sync": (c_int, [c_vp]),
    "fd23_set_punch": (c_int, [c_vp, c_vp, c_flt, c_flt, c_flt]),
    "fd23_set_punch_f64": (c_int, [c_vp, c_vp, c_dbl, c_dbl, c_dbl]),
    "fd23_punch_blocks": (c_int, [c_vp, c_int, c_int, c_flt, c_flt, c_flt, c_vp, c_vp]),
    "fd23_punch_blocks_f64": (c_int, [c_vp, c_int, c_int, c_dbl, c_dbl, c_dbl, c_vp, c_vp]),
    "fd23_get_seismograms": (c_int, [c_vp, c_int, c_int, c_vp, c_int]),
    "fd23_get_field": (c_int, [c_vp, c_int, c_int, c_vp]),
    "fd23_set_field": (c_int, [c_vp, c_int, c_int, c_vp]),
    "fd23_state_bytes": (c_sz, [c_vp]),
    "fd23_get_state": (c_int, [c_vp, c_int, c_vp, c_sz]),
    "fd23_set_state": (c_int, [c_vp, c_int, c_vp, c_sz]),
    "fd23_box_configure": (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_dbl, c_dbl, c_dbl,
                                   c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl]),
    "fd23_box_configure_frames": (c_int, [c_vp, c_int, c_int, c_int, c_vp, c_int, c_dbl, c_dbl, c_dbl, c_dbl,
                                          c_dbl]),
    "fd23_box_forced_cells": (c_int, [c_vp]),
    "fd23_box_get_residual": (c_int, [c_vp, c_int, c_vp]),
    "fd23_set_tile": (c_int, [c_vp, c_int]),
    "fd23_get_tile": (c_int, [c_vp]),
    "fd23_num_tiles": (c_int, []),
    "fd23_tile_shape": (c_int, [c_int, P_int, P_int, P_int, P_int]),
    "fd23_tile_kind": (c_int, [c_int]),
    "fd23_tile_split": (c_int, [c_int]),
    "fd23_autotune": (c_int, [c_vp, c_int, P_dbl]),
    "fd23_set_chains": (c_int, [c_vp, c_int]),
    "fd23_get_chains": (c_int, [c_vp]),
    "fd23_bytes_per_step": (c_dbl, [c_vp]),
    "fd23_device_copy": (c_int, [c_vp, c_vp, c_sz, c_int, c_vp]),
    "fd23_device_stream_tiles": (c_int, [c_vp, c_int, c_vp, c_int, c_int, c_int, c_vp]),
    "fd23_time_passes": (c_int, [c_vp, c_int, P_dbl, P_dbl]),
    "fd23_cpml_profile": (c_int, [c_int, c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_vp, c_vp]),
    "fd23_cpml_pack": (c_int, [c_int, c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_vp]),
    "fd23_cpml_pack_len": (c_sz, [c_int, c_int, c_int]),
    "fd23_psv_material": (c_int, [c_int, c_int, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                  P_dbl]),
    "fd23_psv_material_f64": (c_int, [c_int, c_int, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                      c_vp, P_dbl]),
    "fd23_sh_material": (c_int, [c_int, c_int, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp, P_dbl]),
    "fd23_psv_moment": (c_int, [c_int, c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_vp]),
    "fd23_sh_moment": (c_int, [c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_vp]),
    "fd23_stf_gaussian": (c_int, [c_int, c_vp, c_int, P_int]),
    "fd23_stf_gaussian_f64": (c_int, [c_int, c_vp, c_int, P_int]),
    "fd23_stf_trapezoid_f64": (c_int, [c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_int, P_int]),
    "fd23_psv_moment_f64": (c_int, [c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_vp]),
    "fd23_model_read_f64": (c_int, [ctypes.c_char_p, c_int, c_int, c_dbl, c_vp, c_vp, c_vp]),
    "fd23_isis_write_f64": (c_int, [ctypes.c_char_p, c_vp, c_int, c_int, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl,
                                    c_int, c_int, c_int, c_dbl, c_dbl, c_int]),
    "fd23_stf_trapezoid": (c_int, [c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_int, P_int]),
    "fd23_stf_ricker": (c_int, [c_int, c_dbl, c_dbl, c_dbl, c_vp]),
    "fd23_model_read": (c_int, [ctypes.c_char_p, c_int, c_int, c_dbl, c_vp, c_vp, c_vp]),
    "fd23_isis_write": (c_int, [ctypes.c_char_p, c_vp, c_int, c_int, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl,
                                c_int, c_int, c_int]),
    "fd23_line_receivers": (c_int, [c_int, c_int, c_int, c_int, c_int, c_vp, c_vp]),
    "fd23_station_read": (c_int, [ctypes.c_char_p, c_vp, c_vp, c_int, P_int]),
    "fd23_par_open": (c_int, [c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(c_vp)]),
    "fd23_par_close": (None, [c_vp]),
    "fd23_par_int": (c_int, [c_vp, ctypes.c_char_p, P_int]),
    "fd23_par_float": (c_int, [c_vp, ctypes.c_char_p, P_flt]),
    "fd23_par_double": (c_int, [c_vp, ctypes.c_char_p, P_dbl]),
    "fd23_par_string": (c_int, [c_vp, ctypes.c_char_p, ctypes.c_char_p, c_int]),
    "fd23_set_scales": (c_int, [c_vp, c_vp, c_int]),
    "fd23_server_create": (c_int, [ctypes.POINTER(ServerDesc), ctypes.POINTER(c_vp)]),
    "fd23_server_destroy": (c_int, [c_vp]),
    "fd23_server_submit": (c_int, [c_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_longlong)]),
    "fd23_server_wait": (c_int, [c_vp, ctypes.c_longlong, c_dbl, P_int]),
    "fd23_server_result_bytes": (c_sz, [c_vp, ctypes.c_longlong]),
    "fd23_server_result": (c_int, [c_vp, ctypes.c_longlong, c_vp, c_sz, P_dbl, P_dbl]),
    "fd23_server_stats_get": (c_int, [c_vp, ctypes.POINTER(ServerStats)]),
    "fd23_get_scales": (c_int, [c_vp, c_vp, c_vp, c_int]),
}

_LIB = None


class Fd23Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fd23 error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libfd23.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libfd23.so is not built (%s); run `python -c 'import __graft_entry__ as g; g.build()'` "
            "-- there is no CPU fallback for the fd23 HIP path" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _LIB = lib
    return lib


def check(code):
    if code != OK:
        msg = load().fd23_last_error()
        raise Fd23Error(code, msg.decode() if msg else "")
    return code
