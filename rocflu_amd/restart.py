"""Fluid dumps (restart files) of the reference's RFLUDUMP layout, host arrays.

The codec is the library's (``rflu_dump_write`` / ``rflu_dump_read`` /
``rflu_dump_name`` in include/rocflu_amd.h; modflu/RFLU_ModRestartData.F90).  A region
on the GPU writes and reads its state with ``GpuRegion.write_dump`` / ``read_dump``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _native as nat


def dump_name(out_dir: str, step: int, region: int) -> str:
    """restart_data/fluid_<step:08d>_<region:05d>.dat under out_dir (FluidDumpName,
    modflu/RFLU_ModRestartData.F90:258-266)."""
    buf = ctypes.create_string_buffer(4096)
    nat.check(nat.lib().rflu_dump_name(buf, len(buf), os.fsencode(out_dir), int(step),
                                       int(region)), "rflu_dump_name")
    return os.fsdecode(buf.value)


def write_dump(path: str, cv, step: int, time: float, region: int = 0):
    """Writes cv (n_cells, n_cv) as a version-5 pure-form dump (RFLU_WriteRestartData)."""
    cv = np.ascontiguousarray(cv, dtype=np.float64)
    if cv.ndim != 2:
        raise ValueError("cv must be (n_cells, n_cv)")
    nat.check(nat.lib().rflu_dump_write(os.fsencode(path), int(step), float(time), int(region),
                                        cv.shape[0], cv.shape[1], nat.ptr(cv, ctypes.c_double)),
              "rflu_dump_write")


def read_dump(path: str, n_cells: int, n_cv: int = 5):
    """Reads a dump of a gas-only region with n_cells cells and n_cv variables, with the
    reference reader's checks (RFLU_ReadRestartData); returns (cv, step, time, region)."""
    cv = np.empty((int(n_cells), int(n_cv)))
    step, time, region = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int32()
    nat.check(nat.lib().rflu_dump_read(os.fsencode(path), int(n_cells), int(n_cv),
                                       ctypes.byref(step), ctypes.byref(time), ctypes.byref(region),
                                       nat.ptr(cv, ctypes.c_double)),
              "rflu_dump_read")
    return cv, step.value, time.value, region.value
