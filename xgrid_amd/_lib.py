"""ctypes binding of libxgrid_hip.so (declarations in include/xgrid_hip.h).

The library is built in-tree (``python -m xgrid_amd.build_native``).  There is
no fallback: when it is missing, or no GPU is present, using the device path
raises.
"""
from __future__ import annotations

import ctypes
import os
import threading
from ctypes import (POINTER, Structure, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxgrid_hip.so")

XG_MAX_REFS = 16
XG_MAX_HAZ = 8

MODE_INPLACE, MODE_PINGPONG, MODE_TWOPHASE = 0, 1, 2
DOM_BOX, DOM_LIST, DOM_MASKED, DOM_EMPTY = 0, 1, 2, 3
T_F32, T_F64 = 1, 2


class BackendError(RuntimeError):
    """A call into libxgrid_hip failed; the message is the library's."""


class StmtDesc(Structure):
    _fields_ = [
        ("store_arg", c_int32), ("klass", c_int32), ("mode", c_int32), ("dims", c_int32),
        ("elem_bytes", c_int32), ("nrefs", c_int32),
        ("ref_arg", c_int32 * XG_MAX_REFS), ("ref_level", c_int32 * XG_MAX_REFS),
        ("store_off", c_int32 * 3), ("nhazard", c_int32), ("hazard", (c_int32 * 3) * XG_MAX_HAZ),
        ("param_bytes", c_int32), ("overstep", c_int32), ("group", c_int32), ("group_pos", c_int32),
        ("pair_fwd", c_int32), ("tile3_cols", c_int32),
        ("fn_main", c_void_p), ("fn_pingpong", c_void_p), ("fn_phase2", c_void_p),
        ("fn_main_b", c_void_p), ("fn_pingpong_b", c_void_p), ("fn_phase2_b", c_void_p),
        ("fn_pair", c_void_p),
    ]


class GroupDesc(Structure):
    _fields_ = [("first", c_int32), ("count", c_int32), ("fn", c_void_p), ("fn_b", c_void_p)]


XG_LOOP_MAX = 8


class LoopDesc(Structure):
    _fields_ = [("nstmts", c_int32), ("sids", c_int32 * XG_LOOP_MAX), ("grid_arg", c_int32),
                ("core0", c_int32), ("core1", c_int32), ("halo", c_int32), ("tmax", c_int32),
                ("threads", c_int32), ("fn", c_void_p), ("fn_check", c_void_p), ("fn_b", c_void_p),
                ("fn_check_b", c_void_p), ("alt", c_int32), ("alt_pad", c_int32), ("alt_max_cells", c_int64)]


class Api(Structure):
    _fields_ = [("begin", c_void_p), ("stmt", c_void_p), ("end", c_void_p), ("shape", c_void_p),
                ("loop", c_void_p)]


_SIGS = {
    "xg_last_error": ([], c_char_p),
    "xg_init": ([c_int], c_int),
    "xg_initialized": ([], c_int),
    "xg_device_name": ([c_char_p, c_int64, POINTER(c_int), POINTER(c_int64)], c_int),
    "xg_synchronize": ([], c_int),
    "xg_mem_info": ([POINTER(c_int64), POINTER(c_int64)], c_int),
    "xg_rtc_compile": ([c_char_p, c_char_p, POINTER(c_char_p), c_int32, POINTER(c_char_p),
                        POINTER(c_char_p), c_int32, c_char_p], c_int),
    "xg_module_load": ([c_char_p, POINTER(c_void_p)], c_int),
    "xg_module_function": ([c_void_p, c_char_p, POINTER(c_void_p)], c_int),
    "xg_module_unload": ([c_void_p], c_int),
    "xg_grid_create": ([c_int32, POINTER(c_int32), c_int32, POINTER(c_void_p)], c_int),
    "xg_grid_destroy": ([c_void_p], c_int),
    "xg_grid_set_depth": ([c_void_p, c_int32], c_int),
    "xg_grid_depth": ([c_void_p], c_int32),
    "xg_grid_tick": ([c_void_p], c_int),
    "xg_grid_upload": ([c_void_p, c_int32, c_void_p], c_int),
    "xg_grid_download": ([c_void_p, c_int32, c_void_p], c_int),
    "xg_grid_download_after": ([c_void_p, c_int32, c_void_p, c_void_p], c_int),
    "xg_grid_set_mask": ([c_void_p, c_void_p], c_int),
    "xg_grid_class_info": ([c_void_p, c_int32, POINTER(c_int32), POINTER(c_int64), POINTER(c_int32),
                            POINTER(c_int32)], c_int),
    "xg_grid_level_ptr": ([c_void_p, c_int32], c_void_p),
    "xg_kernel_create": ([c_int32, POINTER(StmtDesc), c_int32, POINTER(GroupDesc), c_int32, POINTER(LoopDesc),
                          c_int32, POINTER(c_void_p)], c_int),
    "xg_kernel_destroy": ([c_void_p], c_int),
    "xg_kernel_status": ([c_void_p], c_int),
    "xg_kernel_set_options": ([c_void_p, c_int32, c_int32, c_int32], c_int),
    "xg_kernel_loop_info": ([c_void_p, c_int32, POINTER(c_int64), POINTER(c_int64), POINTER(c_int32)], c_int),
    "xg_kernel_time_loop": ([c_void_p, c_int32, c_int32, c_int32, POINTER(c_double)], c_int),
    "xg_kernel_counters": ([c_void_p, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64),
                            POINTER(c_int64)], c_int),
    "xg_kernel_time_stmt": ([c_void_p, c_int32, c_int32], c_int),
    "xg_kernel_stmt_time": ([c_void_p, c_int32, POINTER(c_double), POINTER(c_int64)], c_int),
    "xg_kernel_stmt_counts": ([c_void_p, POINTER(c_int64), c_int32, c_int32], c_int),
    "xg_kernel_time_isolated": ([c_void_p, c_int32, c_int32, POINTER(c_double)], c_int),
    "xg_api_table": ([], c_void_p),
    "xg_copy_probe": ([c_int64, c_int32, POINTER(c_double), POINTER(c_double)], c_int),
    "xg_event_record": ([POINTER(c_void_p)], c_int),
    "xg_event_elapsed": ([c_void_p, c_void_p, POINTER(c_double)], c_int),
    "xg_event_destroy": ([c_void_p], c_int),
    "xg_batch_create": ([c_void_p, c_int32, POINTER(c_void_p)], c_int),
    "xg_batch_destroy": ([c_void_p], c_int),
    "xg_batch_add": ([c_void_p, c_void_p, c_void_p, c_int64, POINTER(c_void_p), c_int32, POINTER(c_int64)], c_int),
    "xg_batch_remove": ([c_void_p, c_int64], c_int),
    "xg_batch_size": ([c_void_p], c_int),
    "xg_batch_set_options": ([c_void_p, c_int32, c_int64], c_int),
    "xg_batch_step": ([c_void_p, c_int32], c_int),
    "xg_batch_counters": ([c_void_p, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64), POINTER(c_int64),
                           POINTER(c_int64)], c_int),
    "xg_grid_norm_diff": ([c_void_p, c_int32, c_int32, c_void_p, c_int32, c_int32, c_int32, POINTER(c_double),
                           POINTER(c_double), POINTER(c_int64)], c_int),
    "xg_grid_norm_diff_noalloc": ([c_void_p, c_int32, c_int32, c_void_p, c_int32, c_int32, c_int32, POINTER(c_double),
                                   POINTER(c_double), POINTER(c_int64)], c_int),
    "xg_grid_poisson_residual": ([c_void_p, c_int32, c_int32, c_void_p, c_int32, c_int32, c_int32, POINTER(c_double),
                                  c_double, POINTER(c_double), POINTER(c_double)], c_int),
}

EXPORTED = tuple(_SIGS)

_lib = None
_load_lock = threading.Lock()


def load():
    """Load the library (once) and declare every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    with _load_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise BackendError(f"{LIB_PATH} is missing: build it with `python -m xgrid_amd.build_native`")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (argtypes, restype) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _lib = lib
    return _lib


def last_error() -> str:
    msg = load().xg_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise BackendError(f"{what}: {last_error()} (code {rc})" if what else f"{last_error()} (code {rc})")


_initialised_device = None


def ensure_device(device: int = 0):
    """Initialise the backend on `device` (idempotent)."""
    global _initialised_device
    lib = load()
    if _initialised_device is None:
        check(lib.xg_init(int(device)), "xg_init")
        _initialised_device = int(device)
    return lib
