"""ctypes binding of libmlod.so (the C ABI declared in include/mlod.h).

There is no CPU fallback: if the HIP library is missing, importing the projection classes
raises.  Build it with ``python -m mlodsurfaces_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# MLOD_LIB_PATH selects another build of the same library (e.g. an alternative kernel variant).
LIB_PATH = os.environ.get("MLOD_LIB_PATH") or os.path.join(HERE, "libmlod.so")

MLOD_OK = 0
MAX_SIGMAS = 32

KERNEL_GAUSSIAN, KERNEL_SINGULAR, KERNEL_GAUSSIAN_MULTIPLE = 0, 1, 2
SEM_REF_GPU, SEM_REF_CPU = 0, 1
PREC_FP32, PREC_FP32C, PREC_FP64 = 0, 1, 2
MODE_MLOD, MODE_GLOBAL, MODE_KNN, MODE_BALL = 0, 1, 2, 3
ERR_INVALID, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED, ERR_STATE = -1, -2, -3, -4, -5

# every function the header declares: name -> (restype, argtypes)
_vp = C.c_void_p
_f = C.POINTER(C.c_float)
_i = C.POINTER(C.c_int32)
_u = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)


class KernelC(C.Structure):
    _fields_ = [("type", C.c_int32), ("sigma", C.c_float), ("exponent", C.c_float), ("constant_shift", C.c_float),
                ("n_sigmas", C.c_int32), ("sigmas", C.c_float * MAX_SIGMAS)]


class BuildOptsC(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("max_pts_per_leaf", C.c_uint32), ("tie_policy", C.c_int32),
                ("use_root_box", C.c_uint32), ("root_box", C.c_float * 6)]


class TreeInfoC(C.Structure):
    _fields_ = [("n_points", C.c_uint32), ("n_nodes", C.c_uint32), ("depth", C.c_uint32),
                ("level_offset", C.c_uint32 * 14), ("bbox", C.c_float * 6), ("radius", C.c_float),
                ("ties", C.c_uint64), ("device_bytes", C.c_uint64)]


class ServerDoneC(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("latency_ms", C.c_double), ("queue_ms", C.c_double),
                ("completed_s", C.c_double), ("status", C.c_int32), ("reserved", C.c_int32)]


SIGNATURES = {
    "mlod_last_error": (C.c_char_p, []),
    "mlod_tree_create": (C.c_int, [_f, _f, C.c_uint32, C.POINTER(BuildOptsC), _vp, C.POINTER(_vp)]),
    "mlod_tree_create_device": (C.c_int, [_vp, _vp, C.c_uint32, C.POINTER(BuildOptsC), _vp, C.POINTER(_vp)]),
    "mlod_tree_destroy": (C.c_int, [_vp]),
    "mlod_tree_get_info": (C.c_int, [_vp, C.POINTER(TreeInfoC)]),
    "mlod_set_kernel": (C.c_int, [_vp, C.POINTER(KernelC)]),
    "mlod_set_params": (C.c_int, [_vp, C.c_float, C.c_uint32, C.c_int32, C.c_int32]),
    "mlod_set_mode": (C.c_int, [_vp, C.c_int32, C.c_uint32, C.c_float]),
    "mlod_set_traversal": (C.c_int, [_vp, C.c_int32]),
    "mlod_queries_upload": (C.c_int, [_vp, _f, C.c_uint32, _vp]),
    "mlod_queries_set_device": (C.c_int, [_vp, _vp, C.c_uint32, _vp]),
    "mlod_queries_erase": (C.c_int, [_vp]),
    "mlod_project": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_float, _f, _f, _vp]),
    "mlod_queries_download": (C.c_int, [_vp, C.c_uint32, _f, _f, _vp]),
    "mlod_tree_dump": (C.c_int, [_vp, _i, _f, _f, _f, _i, _i, _i, _i, _u, _u, _f, _f]),
    "mlod_count_visits": (C.c_int, [_vp, C.c_uint32, _u64, _vp]),
    "mlod_server_create": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "mlod_server_destroy": (C.c_int, [_vp]),
    "mlod_server_add_tree": (C.c_int, [_vp, _vp, _u]),
    "mlod_server_submit": (C.c_int, [_vp, C.c_uint32, C.POINTER(KernelC), C.c_float, C.c_uint32, C.c_uint32, C.c_float,
                                     _vp, C.c_uint32, _vp, _vp, _u64]),
    "mlod_server_poll": (C.c_int, [_vp, C.POINTER(ServerDoneC), C.c_uint32, _u, C.c_double]),
    "mlod_server_stats": (C.c_int, [_vp, _u64]),
    "mlod_server_set_budget": (C.c_int, [_vp, C.c_uint64]),
    "mlod_server_memory": (C.c_int, [_vp, _u64]),
    "mlod_server_cancel": (C.c_int, [_vp, C.c_uint64]),
    "mlod_project_points": (C.c_int, [_vp, _f, C.c_uint32, C.c_uint32, C.c_float, _f, _f, _vp]),
    "mlod_project_points_noalloc": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, C.c_float, _vp, _vp, _vp]),
    "mlod_queries_reorganize": (C.c_int, [_vp, _u, C.c_uint32, C.c_uint32, _vp]),
    "mlod_queries_fit_kinds": (C.c_int, [_vp, C.c_uint32, _vp, _vp]),
    "mlod_generator_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mlod_generator_destroy": (C.c_int, [_vp]),
    "mlod_generate": (C.c_int, [_vp, C.c_uint32, C.c_int32, _vp]),
    "mlod_generate_from_points": (C.c_int, [_vp, _vp, C.c_uint32, C.c_int32, C.c_uint32, _vp, _vp]),
    "mlod_generator_info": (C.c_int, [_vp, _u64]),
    "mlod_generator_cells": (C.c_int, [_vp, C.c_int32, _vp, _vp]),
    "mlod_generator_centres": (C.c_int, [_vp, _vp, C.POINTER(_vp), _vp]),
    "mlod_dc_create": (C.c_int, [C.POINTER(_vp)]),
    "mlod_dc_destroy": (C.c_int, [_vp]),
    "mlod_dc_build": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, C.c_int32, _vp]),
    "mlod_dc_build_from_points": (C.c_int, [_vp, _vp, C.c_uint32, _vp]),
    "mlod_dc_info": (C.c_int, [_vp, _vp]),
    "mlod_dc_nodes": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mlod_dc_corners": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mlod_dc_project_corners": (C.c_int, [_vp, _vp, _vp]),
    "mlod_dc_set_corner_values": (C.c_int, [_vp, _vp, _vp]),
    "mlod_dc_corner_values": (C.c_int, [_vp, _vp, _vp]),
    "mlod_dc_extract": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _vp]),
    "mlod_dc_mesh": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "mlod_dc_refine": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), _vp]),
    "mlod_dc_refined_mesh": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mlod_mesh_smooth": (C.c_int, [_vp, C.c_uint32, _vp, C.c_uint32, _vp, _vp]),
    "mlod_mesh_subdivide": (C.c_int, [_vp, C.c_uint32, _vp, C.c_uint32, _vp, C.c_uint32, _vp, C.POINTER(C.c_uint32), _vp]),
}

_lib = None


class MlodError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmlod error {code}: {msg}")
        self.code = code


def lib():
    """Loads libmlod.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m mlodsurfaces_amd.build`")
        L = C.CDLL(LIB_PATH)
        override = bool(os.environ.get("MLOD_LIB_PATH"))
        for name, (res, args) in SIGNATURES.items():
            if override and not hasattr(L, name):
                continue  # an older library given explicitly: its missing entries fail when called
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != MLOD_OK:
        msg = lib().mlod_last_error()
        raise MlodError(code, msg.decode() if msg else "")
    return code
