"""ctypes binding of libflowhip.so (include/flowhip.h).

Loads the in-tree library and fails loudly when it is missing: there is no
CPU fallback anywhere in flowlab_amd.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflowhip.so")

FH_OK, FH_ERR_BAD_ARG, FH_ERR_BUSY, FH_ERR_DEVICE_LOST, FH_ERR_NO_DEVICE, FH_ERR_HIP = range(6)
SEMI_LAGRANGIAN, MACCORMACK = 0, 1
SOLVER_RBSOR, SOLVER_MG, SOLVER_MG_DC, SOLVER_MG_FCG = 1, 2, 3, 4
PHASES = ("advect", "forces", "divergence", "solve", "project", "dye", "total")

SCHEMES = {"semiLagrangian": SEMI_LAGRANGIAN, "macCormack": MACCORMACK}
SOLVERS = {"rbsor": SOLVER_RBSOR, "mg": SOLVER_MG, "mg-dc": SOLVER_MG_DC,
           "mg-fcg": SOLVER_MG_FCG}

# Every exported symbol of include/flowhip.h: (name, restype, argtypes).
_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_i, _d = C.c_int, C.c_double

SIGNATURES = [
    ("fh_last_error", C.c_char_p, []),
    ("fh_device_count", _i, []),
    ("fh_create", _vp, [_i, _i, _i, _d]),
    ("fh_destroy", None, [_vp]),
    ("fh_upload", _i, [_vp, _dp, _dp, _dp, _u8p, _dp]),
    ("fh_set_labels", _i, [_vp, _u8p]),
    ("fh_set_dye_patch", _i, [_vp, _i, _i, _i, _i, _fp]),
    ("fh_set_buoyancy", _i, [_vp, _d]),
    ("fh_set_confinement", _i, [_vp, _d]),
    ("fh_set_stir", _i, [_vp, _i, _d, _d, _d, _d, _d]),
    ("fh_step", _i, [_vp, _i, _d, _d, _d, _dp, _i, _i, _d, _d, _ip]),
    ("fh_read_velocity", _i, [_vp, _dp, _dp]),
    ("fh_read_dye", _i, [_vp, _dp]),
    ("fh_read_pressure", _i, [_vp, _dp]),
    ("fh_read_divergence", _i, [_vp, _dp]),
    ("fh_max_face_speed", _i, [_vp, _dp]),
    ("fh_solve_pressure", _i, [_vp, _i, _dp, _dp, _u8p, _d, _i, _d, _d, _ip]),
    ("fh_advect_velocity", _i, [_vp, _i, _dp, _dp, _dp, _dp, _u8p, _d]),
    ("fh_advect_dye", _i, [_vp, _i, _dp, _dp, _dp, _dp, _u8p, _d]),
    ("fh_phase_advect_velocity", _i, [_vp, _i, _d]),
    ("fh_phase_forces", _i, [_vp, _d]),
    ("fh_phase_divergence", _i, [_vp]),
    ("fh_phase_solve", _i, [_vp, _i, _d, _i, _d, _d, _ip]),
    ("fh_phase_project", _i, [_vp, _d]),
    ("fh_phase_dye", _i, [_vp, _i, _d, _dp]),
    ("fh_solve_history", _i, [_vp, _dp, _i]),
    ("fh_set_timing", _i, [_vp, _i]),
    ("fh_get_timings", _i, [_vp, _dp]),
    ("fh_synchronize", _i, [_vp]),
    ("fh_trace_kernel", _i, [_vp, _i]),
    ("fh_trace_stats", _i, [_vp, _ip, _dp]),
    ("fh_trace_overhead", _i, [_vp, _i, _dp]),
    ("fh_num_tile_shapes", _i, []),
    ("fh_tile_shape_name", C.c_char_p, [_i]),
    ("fh_set_tile", _i, [_vp, _i, _i]),
    ("fh_get_tiles", _i, [_vp, _ip, _ip]),
    ("fh_tune_tiles", _i, [_vp, _i, _fp]),
    ("fh_set_coarse_group", _i, [_vp, _i]),
    ("fh_set_advect_staging", _i, [_vp, _i]),
    ("fh_set_dye_one_pass", _i, [_vp, _i]),
    ("fh_set_faces_one_pass", _i, [_vp, _i]),
    ("fh_coarse_group_threads", _i, [_vp]),
    ("fh_noalloc_last_error", C.c_char_p, []),
    ("fh_noalloc_max_cells", C.c_longlong, []),
    ("fh_solve_pressure_rbsor_noalloc", _i, [_vp, _vp, _vp, _vp, _i, _i, _d, _i, _d, _d, _ip]),
    ("fh_unet_create", _vp, [C.c_ulonglong]),
    ("fh_unet_destroy", None, [_vp]),
    ("fh_unet_last_error", C.c_char_p, []),
    ("fh_unet_num_params", _i, []),
    ("fh_unet_get_params", _i, [_vp, _fp, _i]),
    ("fh_unet_set_params", _i, [_vp, _fp, _i]),
    ("fh_unet_coarse_grid", _i, [_i, _i, _ip, _ip]),
    ("fh_unet_set_scale", _i, [_vp, _d]),
    ("fh_unet_get_scale", _i, [_vp, _dp]),
    ("fh_unet_forward", _i, [_vp, _i, _i, _i, _fp, _fp]),
    ("fh_unet_backward", _i, [_vp, _i, _i, _i, _fp, _fp, _fp]),
    ("fh_unet_dataset_reserve", _i, [_vp, _i, _i, _i]),
    ("fh_unet_dataset_size", _i, [_vp]),
    ("fh_unet_dataset_get", _i, [_vp, _i, _fp, _fp, _fp]),
    ("fh_unet_calibrate", _i, [_vp, _i]),
    ("fh_unet_train", _i, [_vp, _i, _i, _d, C.c_ulonglong, _dp]),
    ("fh_set_capture", _i, [_vp, _vp]),
    ("fh_last_solve_ms", _i, [_vp, _dp]),
    ("fh_set_warm_start", _i, [_vp, _vp]),
    ("fh_set_warm_extrapolation", _i, [_vp, _i]),
    ("fh_set_fused_update", _i, [_vp, _i]),
    ("fh_batch_last_error", C.c_char_p, []),
    ("fh_batch_create", _vp, [_i, _ip, _ip, _dp]),
    ("fh_batch_destroy", None, [_vp]),
    ("fh_batch_upload", _i, [_vp, _i, _dp, _dp, _dp, _u8p, _dp]),
    ("fh_batch_set_forces", _i, [_vp, _i, _d, _d, _dp]),
    ("fh_batch_step", _i, [_vp, _i, _d, _d, _d, _d, _i]),
    ("fh_batch_results", _i, [_vp, _dp, _ip, _dp]),
    ("fh_batch_set_warm_extrapolation", _i, [_vp, _i]),
    ("fh_batch_read", _i, [_vp, _i, _dp, _dp, _dp, _dp]),
    ("fh_batch_last_step_ms", _i, [_vp, _fp]),
    ("fh_batch_cells", C.c_longlong, [_vp]),
    ("fh_batch_num_tiles", _i, []),
    ("fh_batch_tile_name", C.c_char_p, [_i]),
    ("fh_batch_set_tile", _i, [_vp, _i, _i]),
    ("fh_batch_get_tiles", _i, [_vp, _ip, _ip]),
    ("fh_batch_tune", _i, [_vp, _i, _fp]),
]

KERNELS = {"rb_sweep_fine": 1, "advect_faces": 2, "maccormack_faces": 3, "advect_dye": 4,
           "maccormack_dye": 5, "dc_update": 6, "restrict_fine": 7,
           "smooth_restrict_fine": 8, "prolong_smooth_fine": 9, "coarse_vcycle": 10,
           "fcg_dots": 11, "fcg_update": 12, "divergence": 13, "forces": 14, "project": 15}


class HipError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
            " (flowlab_amd has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def last_error() -> str:
    return (lib().fh_last_error() or b"").decode()


def check(status: int) -> None:
    """Status -> exception, with the reference's messages: a re-entrant call
    says '<Class>.<method>() is not re-entrant' and a lost device says
    'GPU device lost' (src/gpu/stepGpu.ts:402-403)."""
    if status == FH_OK:
        return
    msg = last_error()
    if status == FH_ERR_DEVICE_LOST:
        msg = "GPU device lost"
    raise HipError(status, msg)


def check_batch(status: int) -> None:
    """check() for the fh_batch_* entry points (their own error string)."""
    if status == FH_OK:
        return
    msg = (lib().fh_batch_last_error() or b"").decode()
    if status == FH_ERR_DEVICE_LOST:
        msg = "GPU device lost"
    raise HipError(status, msg)


def f64(a: np.ndarray) -> np.ndarray:
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 array")
    return a


def _sized(a, n, what):
    """The library reads or writes exactly n elements through the pointer:
    a shorter (or longer) array is a caller error, raised before any call."""
    if n is not None and a.size != n:
        raise ValueError(f"{what}: expected {n} elements, got {a.size}")
    return a


def dptr(a, n=None, what="float64 field"):
    return None if a is None else _sized(f64(a), n, what).ctypes.data_as(_dp)


def u8ptr(a, n=None, what="label"):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.uint8 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous uint8 label array")
    return _sized(a, n, what).ctypes.data_as(_u8p)


def fptr(a, n=None, what="float32 array"):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float32 array")
    return _sized(a, n, what).ctypes.data_as(_fp)
