"""ctypes bindings of ``librocflu_amd.so`` (the C-ABI declared in include/rocflu_amd.h).

Loading fails loudly: there is no CPU fallback behind these calls.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librocflu_amd.so")

RFLU_OK = 0
RFLU_ERR_INVALID_VALUE = 1
RFLU_ERR_NEGATIVE_POSDEF = 2
RFLU_ERR_ARG = 3
RFLU_ERR_HIP = 4
RFLU_ERR_NOMEM = 5
RFLU_ERR_HANDLE = 6
RFLU_ERR_FILE = 7
RFLU_ERR_DUMP = 8

ERROR_NAMES = {
    1: "ERR_INVALID_VALUE", 2: "ERR_NEGATIVE_POSDEF", 3: "ERR_ARG", 4: "ERR_HIP",
    5: "ERR_NOMEM", 6: "ERR_HANDLE", 7: "ERR_FILE", 8: "ERR_DUMP",
}

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class RegionDesc(ctypes.Structure):
    """Mirror of ``rflu_region_desc``."""
    _fields_ = [
        ("index_base", ctypes.c_int32), ("n_cells", ctypes.c_int32),
        ("n_faces", ctypes.c_int32), ("n_patches", ctypes.c_int32),
        ("f2c", _i32p), ("fn", _f64p), ("fc", _f64p), ("cofg", _f64p), ("vol", _f64p),
        ("bpatch_ptr", _i32p), ("bf2c", _i32p), ("bfn", _f64p), ("bfc", _f64p),
        ("patch_bc", _i32p), ("patch_opt", _i32p), ("patch_pout", _f64p),
        ("patch_order", _i32p),
        ("c2cs_ptr", _i32p), ("c2cs_idx", _i32p), ("xyz_moms", _f64p),
        ("gamma", ctypes.c_double), ("cp", ctypes.c_double),
        ("discr", ctypes.c_int32), ("order", ctypes.c_int32), ("reconst", ctypes.c_int32),
        ("rk", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("cfl", ctypes.c_double), ("dt_imposed", ctypes.c_double),
        ("dt_min_limit", ctypes.c_double), ("eps_entr", ctypes.c_double),
        ("diss_fact", ctypes.c_double),
        ("patch_vals", _f64p),
        ("c2cs_bptr", _i32p), ("c2cs_bface", _i32p),
        ("creconst_weight", ctypes.c_double), ("creconst", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("patch_opt2", _i32p),
        ("n_virtual", ctypes.c_int32),
        ("pad2_", ctypes.c_int32),
    ]


class GpuStats(ctypes.Structure):
    """Mirror of ``rflu_gpu_stats``."""
    _fields_ = [
        ("n_invalid", ctypes.c_int64), ("n_negative", ctypes.c_int64),
        ("n_bounds_fix", ctypes.c_int64), ("n_pos_fix", ctypes.c_int64),
        ("mass_in", ctypes.c_double), ("mass_out", ctypes.c_double),
        ("time", ctypes.c_double), ("dt", ctypes.c_double), ("steps", ctypes.c_int64),
        ("finished", ctypes.c_int32), ("pad", ctypes.c_int32),
    ]


# name -> (restype, argtypes); every symbol include/rocflu_amd.h declares
SIGNATURES = {
    "rflu_gpu_device_count": (ctypes.c_int, []),
    "rflu_gpu_create_region": (ctypes.c_int, [ctypes.POINTER(RegionDesc), _i32p]),
    "rflu_gpu_create_batch": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(RegionDesc), _i32p]),
    "rflu_gpu_destroy_region": (ctypes.c_int, [ctypes.c_int32]),
    "rflu_gpu_upload_state": (ctypes.c_int, [ctypes.c_int32, _f64p, _f64p, ctypes.c_double]),
    "rflu_gpu_download_state": (ctypes.c_int, [ctypes.c_int32, _f64p, _f64p]),
    "rflu_gpu_time_step": (ctypes.c_int, [ctypes.c_int32, ctypes.c_double, _f64p]),
    "rflu_gpu_rk_init": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32]),
    "rflu_gpu_cell_gradients": (ctypes.c_int, [ctypes.c_int32]),
    "rflu_gpu_convective_fluxes": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32]),
    "rflu_gpu_rk_update": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double]),
    "rflu_gpu_set_end_time": (ctypes.c_int, [ctypes.c_int32, ctypes.c_double]),
    "rflu_gpu_set_dependent_vars": (ctypes.c_int, [ctypes.c_int32]),
    "rflu_gpu_after_update": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32]),
    "rflu_gpu_check_status": (ctypes.c_int, [ctypes.c_int32]),
    "rflu_gpu_residual": (ctypes.c_int, [ctypes.c_int32, _f64p]),
    "rflu_gpu_time_steps": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32]),
    "rflu_gpu_set_borders": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _i32p,
                                            _i32p, _i32p, _i32p]),
    "rflu_gpu_exchange": (ctypes.c_int, [ctypes.c_int32, _i32p, ctypes.c_int32]),
    "rflu_gpu_set_border_transform": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _f64p]),
    "rflu_gpu_get_stats": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(GpuStats)]),
    "rflu_gpu_get_case_stats": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(GpuStats)]),
    "rflu_gpu_cell_order": (ctypes.c_int, [ctypes.c_int32, _i32p]),
    "rflu_gpu_cell_volumes": (ctypes.c_int, [ctypes.c_int32, _f64p]),
    "rflu_gpu_patch_faces": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _i32p, _f64p]),
    "rflu_gpu_set_patch_values": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _f64p]),
    "rflu_gpu_set_state_dev": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p]),
    "rflu_gpu_get_state_dev": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p]),
    "rflu_gpu_residual_dev": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p]),
    "rflu_gpu_residual_jvp": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, _f64p, ctypes.c_void_p]),
    "rflu_gpu_residual_jvp_static": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, _f64p, ctypes.c_void_p]),
    "rflu_gpu_residual_vjp": (ctypes.c_int, [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, _f64p]),
    "rflu_gpu_reset_stats": (ctypes.c_int, [ctypes.c_int32]),
    "rflu_gpu_synchronize": (ctypes.c_int, [ctypes.c_int32]),
    "rflu_gpu_set_launch": (ctypes.c_int, [ctypes.c_int32] * 4),
    "rflu_gpu_time_kernel": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f64p]),
    "rflu_gpu_set_profiling": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32]),
    "rflu_gpu_get_kernel_times": (ctypes.c_int, [ctypes.c_int32, _f64p,
                                                 ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]),
    "rflu_gpu_set_option": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "rflu_gpu_get_option": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _i32p]),
    "rflu_dump_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p,
                                      ctypes.c_int64, ctypes.c_int32]),
    "rflu_dump_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_double,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f64p]),
    "rflu_dump_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_int64), _f64p, _i32p, _f64p]),
    "rflu_gpu_write_dump": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p,
                                           ctypes.c_int64, ctypes.c_double, ctypes.c_int32]),
    "rflu_gpu_read_dump": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p,
                                          ctypes.POINTER(ctypes.c_int64), _f64p, _i32p]),
    "rflu_gpu_last_error": (ctypes.c_char_p, []),
}

_LIB = None


class RfluError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where}: {ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Load the HIP library; raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first "
                               "(the GPU path has no CPU fallback)")
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def check(rc: int, where: str) -> None:
    if rc != RFLU_OK:
        raise RfluError(rc, where, lib().rflu_gpu_last_error().decode())


def ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))
