"""Host-side mirror of the reference's region and stage loop over the C-ABI.

``GpuRegion`` owns one region's device copy (the t_region the reference keeps per
MPI rank) and exposes RungeKuttaMP's callees one by one (``rk_init``,
``cell_gradients``, ``convective_fluxes``, ``rk_update``, ``set_dependent_vars``,
``after_update``; libfloflu/RungeKuttaMP.F90:148-508) as well as the fused time loop
(``time_steps``; rocflu/RFLU_TimeStepping.F90:353-640).  Errors come back as
``RfluError`` with the reference's error names (ERR_INVALID_VALUE,
ERR_NEGATIVE_POSDEF).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native as nat
from .mesh import Mesh

DISCR_ROE, DISCR_HLLC, DISCR_AUSMPLUS = 1, 3, 4
RECONST_NONE, RECONST_WENO_XYZ, RECONST_LIM_BARTHJESP, RECONST_LIM_VENKAT = 0, 2, 10, 11
RK4_CLASSICAL, RK3_WRAY, RK1_EULER, RK3_SSP = 1, 2, 3, 4
PREC_FP64, PREC_MIXED, PREC_FP32_GRAD, PREC_FP32_FLUX = 0, 1, 2, 3


def n_stages(rk: int) -> int:
    return {1: 4, 2: 3, 3: 1, 4: 3}[rk]


@dataclass
class Numerics:
    """The t_mixt_input / global fields of the hot path (defaults from
    libflu/RFLU_InitInputValues.F90:182-215 and RFLU_InitGlobal.F90:351-352)."""
    discr: int = DISCR_HLLC
    order: int = 1
    reconst: int = RECONST_NONE
    rk: int = RK3_WRAY
    precision: int = PREC_FP64
    cfl: float = 0.5
    dt_imposed: float = 1.0e-9
    dt_min_limit: float = 0.0
    eps_entr: float = 0.05
    diss_fact: float = 1.0
    gamma: float = 1.4
    cp: float = 1004.5
    patch_order: Optional[Sequence[int]] = None
    # per patch the five BCDAT values of its type: inflow 11 (u, v, w, T, p), inflow 10
    # (ptot, ttot, betah, betav, Mach), farfield (Mach, attack, sideslip, p, T),
    # injection (mass flux, T, -, -, -); zeros elsewhere
    patch_vals: Optional[Sequence[Sequence[float]]] = None
    # per patch BCSWI_INFLOW_FIXED of inflow 10 patches (0 or 1), or BCSWI_FARF_CORR of
    # farfield patches (0 only); None = all 0
    patch_opt2: Optional[Sequence[int]] = None
    # constrained gradients at walls (CRECONSTC, CRECONSTCW; RFLU_InitInputValues.F90:
    # 188-189 defaults 0 and 1.0)
    creconst: int = 0
    creconst_weight: float = 1.0

    def gas_prime(self):
        """(M, R', gamma') of the dependent-variable path (libflu/RFLU_SetGasVars.F90:172-178)."""
        gc0 = self.cp - self.cp / self.gamma
        m = 8314.3 / gc0
        rg = 8314.3 / m
        return m, rg, self.cp / (self.cp - rg)


def region_arrays(mesh: Mesh, num: Numerics) -> dict:
    """Contiguous arrays of one region, kept alive while a descriptor points at them."""
    a = mesh.a
    keep = {
        "f2c": np.ascontiguousarray(a["f2c"], dtype=np.int32).reshape(-1),
        "fn": np.ascontiguousarray(a["fn"], dtype=np.float64).reshape(-1),
        "fc": np.ascontiguousarray(a["fc"], dtype=np.float64).reshape(-1),
        "cofg": np.ascontiguousarray(a["cofg"], dtype=np.float64).reshape(-1),
        "vol": np.ascontiguousarray(a["vol"], dtype=np.float64).reshape(-1),
        "bpatch_ptr": np.ascontiguousarray(a["bpatch_ptr"], dtype=np.int32),
        "bf2c": np.ascontiguousarray(a["bf2c"], dtype=np.int32).reshape(-1),
        "bfn": np.ascontiguousarray(a["bfn"], dtype=np.float64).reshape(-1),
        "bfc": np.ascontiguousarray(a["bfc"], dtype=np.float64).reshape(-1),
        "patch_bc": np.ascontiguousarray(a["patch_bc"], dtype=np.int32),
        "patch_opt": np.ascontiguousarray(a["patch_opt"], dtype=np.int32),
        "patch_pout": np.ascontiguousarray(a["patch_pout"], dtype=np.float64),
        "c2f_ptr": np.ascontiguousarray(a["c2f_ptr"], dtype=np.int32),
        "c2f_idx": np.ascontiguousarray(a["c2f_idx"], dtype=np.int32),
    }
    po = num.patch_order if num.patch_order is not None else [1] * int(mesh.n_patches)
    keep["patch_order"] = np.ascontiguousarray(po, dtype=np.int32)
    pv = np.zeros((int(mesh.n_patches), 5))
    if num.patch_vals is not None:
        pv[:] = np.asarray(num.patch_vals, dtype=np.float64).reshape(-1, 5)
    keep["patch_vals"] = np.ascontiguousarray(pv.reshape(-1))
    o2 = num.patch_opt2 if num.patch_opt2 is not None else [0] * int(mesh.n_patches)
    keep["patch_opt2"] = np.ascontiguousarray(o2, dtype=np.int32)
    if num.order == 2:
        if not mesh.has_stencils:
            raise ValueError("second-order runs need a mesh built with stencils=True")
        keep["c2cs_ptr"] = np.ascontiguousarray(a["c2cs_ptr"], dtype=np.int32)
        keep["c2cs_idx"] = np.ascontiguousarray(a["c2cs_idx"], dtype=np.int32)
        keep["xyz_moms"] = np.ascontiguousarray(a["xyz_moms"], dtype=np.float64).reshape(-1)
        if num.creconst:
            keep["c2cs_bptr"] = np.ascontiguousarray(a["c2cs_bptr"], dtype=np.int32)
            keep["c2cs_bface"] = np.ascontiguousarray(a["c2cs_bface"], dtype=np.int32).reshape(-1)
    return keep


def region_desc(mesh: Mesh, num: Numerics, d: "nat.RegionDesc"):
    """Fills a descriptor for one region; returns the arrays it points at (keep them
    alive until the region is created)."""
    k = region_arrays(mesh, num)
    i32, f64 = ctypes.c_int32, ctypes.c_double
    d.index_base = 0
    d.n_cells = int(mesh.n_cells)
    d.n_virtual = int(getattr(mesh, "n_virtual", 0))
    d.n_faces = int(mesh.n_faces)
    d.n_patches = int(mesh.n_patches)
    d.f2c = nat.ptr(k["f2c"], i32)
    d.fn = nat.ptr(k["fn"], f64)
    d.fc = nat.ptr(k["fc"], f64)
    d.cofg = nat.ptr(k["cofg"], f64)
    d.vol = nat.ptr(k["vol"], f64)
    d.bpatch_ptr = nat.ptr(k["bpatch_ptr"], i32)
    d.bf2c = nat.ptr(k["bf2c"], i32)
    d.bfn = nat.ptr(k["bfn"], f64)
    d.bfc = nat.ptr(k["bfc"], f64)
    d.patch_bc = nat.ptr(k["patch_bc"], i32)
    d.patch_opt = nat.ptr(k["patch_opt"], i32)
    d.patch_pout = nat.ptr(k["patch_pout"], f64)
    d.patch_order = nat.ptr(k["patch_order"], i32)
    if num.order == 2:
        d.c2cs_ptr = nat.ptr(k["c2cs_ptr"], i32)
        d.c2cs_idx = nat.ptr(k["c2cs_idx"], i32)
        d.xyz_moms = nat.ptr(k["xyz_moms"], f64)
    d.gamma, d.cp = num.gamma, num.cp
    d.discr, d.order, d.reconst, d.rk, d.precision = (
        num.discr, num.order, num.reconst, num.rk, num.precision)
    d.cfl, d.dt_imposed, d.dt_min_limit = num.cfl, num.dt_imposed, num.dt_min_limit
    d.eps_entr, d.diss_fact = num.eps_entr, num.diss_fact
    d.patch_vals = nat.ptr(k["patch_vals"], f64)
    if num.order == 2 and num.creconst:
        d.c2cs_bptr = nat.ptr(k["c2cs_bptr"], i32)
        d.c2cs_bface = nat.ptr(k["c2cs_bface"], i32)
    d.creconst_weight = num.creconst_weight
    d.creconst = num.creconst if num.order == 2 else 0
    d.patch_opt2 = nat.ptr(k["patch_opt2"], i32)
    return k


class GpuRegion:
    """One region on the GPU behind the C-ABI."""

    def __init__(self, mesh: Mesh, num: Numerics):
        self.mesh = mesh
        self.num = num
        # actual cells, then virtual cells (a region of a multi-region problem)
        self.n = int(mesh.n_cells) + int(getattr(mesh, "n_virtual", 0))
        L = nat.lib()
        d = nat.RegionDesc()
        keep = region_desc(mesh, num, d)
        h = ctypes.c_int32(0)
        nat.check(L.rflu_gpu_create_region(ctypes.byref(d), ctypes.byref(h)),
                  "rflu_gpu_create_region")
        del keep
        self.handle = h.value
        self._lib = L
        self.k = 1

    def signature(self) -> str:
        m, num = self.mesh, self.num
        s = int(m.c2cs_idx.size) if (num.order == 2 and m.has_stencils) else 0
        return (f"N{m.n_cells}-F{m.n_faces}-B{m.n_bfaces}-S{s}"
                f"-d{num.discr}-o{num.order}-r{num.reconst}-k{num.rk}-p{num.precision}")

    # -- state ---------------------------------------------------------------
    def upload(self, cv: np.ndarray, dv: Optional[np.ndarray] = None, time: float = 0.0):
        cv = np.ascontiguousarray(cv, dtype=np.float64).reshape(self.n, 5)
        dvp = None
        if dv is not None:
            dv = np.ascontiguousarray(dv, dtype=np.float64).reshape(self.n, 3)
            dvp = nat.ptr(dv, ctypes.c_double)
        nat.check(self._lib.rflu_gpu_upload_state(self.handle, nat.ptr(cv, ctypes.c_double),
                                                  dvp, float(time)), "rflu_gpu_upload_state")
        self.time = float(time)

    def download(self):
        cv = np.empty((self.n, 5))
        dv = np.empty((self.n, 3))
        nat.check(self._lib.rflu_gpu_download_state(self.handle, nat.ptr(cv, ctypes.c_double),
                                                    nat.ptr(dv, ctypes.c_double)),
                  "rflu_gpu_download_state")
        return cv, dv

    # -- fluid dumps (modflu/RFLU_ModRestartData.F90) --------------------------
    def write_dump(self, path: str, step: int, time: float, region: int = 0, icase: int = 0):
        """Writes the current state of case icase as the RFLUDUMP file of `region` at
        (step, time) (RFLU_WriteRestartData)."""
        nat.check(self._lib.rflu_gpu_write_dump(self.handle, int(icase), os.fsencode(path),
                                                int(step), float(time), int(region)),
                  "rflu_gpu_write_dump")

    def read_dump(self, path: str, icase: int = 0):
        """Reads an RFLUDUMP file into case icase (state, dependent variables and the
        dump's time, RFLU_ReadRestartData); returns (step, time, region)."""
        step, time, region = ctypes.c_int64(), ctypes.c_double(), ctypes.c_int32()
        nat.check(self._lib.rflu_gpu_read_dump(self.handle, int(icase), os.fsencode(path),
                                               ctypes.byref(step), ctypes.byref(time),
                                               ctypes.byref(region)),
                  "rflu_gpu_read_dump")
        if self.k == 1:
            self.time = time.value
        return step.value, time.value, region.value

    # -- RungeKuttaMP's callees (libfloflu/RungeKuttaMP.F90:216-437) -----------
    def time_step(self, current_time: Optional[float] = None) -> float:
        """dtMin at global%currentTime (the host's time unless given)."""
        t = self.time if current_time is None else float(current_time)
        dt = ctypes.c_double()
        nat.check(self._lib.rflu_gpu_time_step(self.handle, t, ctypes.byref(dt)),
                  "rflu_gpu_time_step")
        return dt.value

    def rk_init(self, stage: int):
        nat.check(self._lib.rflu_gpu_rk_init(self.handle, stage), "rflu_gpu_rk_init")

    def cell_gradients(self):
        nat.check(self._lib.rflu_gpu_cell_gradients(self.handle), "rflu_gpu_cell_gradients")

    def convective_fluxes(self, stage: int):
        nat.check(self._lib.rflu_gpu_convective_fluxes(self.handle, stage),
                  "rflu_gpu_convective_fluxes")

    def rk_update(self, stage: int, dt_min: float):
        nat.check(self._lib.rflu_gpu_rk_update(self.handle, stage, float(dt_min)),
                  "rflu_gpu_rk_update")

    def set_dependent_vars(self):
        nat.check(self._lib.rflu_gpu_set_dependent_vars(self.handle),
                  "rflu_gpu_set_dependent_vars")

    def after_update(self, stage: int):
        """Queues the checks; a failure is reported by the next time_step() or
        check_status() (the state stops where it failed)."""
        nat.check(self._lib.rflu_gpu_after_update(self.handle, stage), "rflu_gpu_after_update")

    def check_status(self):
        """Raises the error of a failed after-update check (waits for the device)."""
        nat.check(self._lib.rflu_gpu_check_status(self.handle), "rflu_gpu_check_status")

    def runge_kutta(self, end_time: Optional[float] = None):
        """One time step through the separate entry points: RFLU_TimeStepping's dt and
        end-time clamp, the stages in RungeKuttaMP's order, and the host-side time
        advance (rocflu/RFLU_TimeStepping.F90:429-640).  Returns (dt, finished)."""
        dt = self.time_step()
        finished = False
        if end_time is not None and self.time + dt > end_time:
            dt = end_time - self.time
            finished = True
        for s in range(1, n_stages(self.num.rk) + 1):
            self.rk_init(s)
            self.cell_gradients()
            self.convective_fluxes(s)
            self.rk_update(s, dt)
            self.set_dependent_vars()
            self.after_update(s)
        self.time = self.time + dt
        return dt, finished

    def set_end_time(self, end_time: float):
        nat.check(self._lib.rflu_gpu_set_end_time(self.handle, float(end_time)),
                  "rflu_gpu_set_end_time")

    # -- fused paths ---------------------------------------------------------------
    def time_steps(self, nsteps: int, graph: bool = False):
        nat.check(self._lib.rflu_gpu_time_steps(self.handle, int(nsteps), 1 if graph else 0),
                  "rflu_gpu_time_steps")
        self.time = self.stats()["time"]

    def residual(self) -> np.ndarray:
        rhs = np.empty((self.n, 5))
        nat.check(self._lib.rflu_gpu_residual(self.handle, nat.ptr(rhs, ctypes.c_double)),
                  "rflu_gpu_residual")
        return rhs

    def stats(self) -> dict:
        s = nat.GpuStats()
        nat.check(self._lib.rflu_gpu_get_stats(self.handle, ctypes.byref(s)), "rflu_gpu_get_stats")
        return {name: getattr(s, name) for name, _ in nat.GpuStats._fields_}

    def reset_stats(self):
        nat.check(self._lib.rflu_gpu_reset_stats(self.handle), "rflu_gpu_reset_stats")

    def synchronize(self):
        nat.check(self._lib.rflu_gpu_synchronize(self.handle), "rflu_gpu_synchronize")

    def set_launch(self, face_block=0, cell_block=0, grad_block=0):
        nat.check(self._lib.rflu_gpu_set_launch(self.handle, face_block, cell_block, grad_block),
                  "rflu_gpu_set_launch")

    def time_kernel(self, which: int, reps: int = 10) -> float:
        ms = ctypes.c_double()
        nat.check(self._lib.rflu_gpu_time_kernel(self.handle, which, reps, ctypes.byref(ms)),
                  "rflu_gpu_time_kernel")
        return ms.value

    OPT_GRAD_LDS, OPT_GRAD_UNROLL, OPT_TILE_CAP, OPT_XCD, OPT_GRAD_WEIGHTS = 1, 2, 3, 4, 5
    OPT_GRAD_TILE = 6
    OPT_FACE_TILE = 7
    OPT_UPDATE_TILE = 8
    OPT_GRAD_FAST = 9
    OPT_PERSISTENT = 10
    OPT_FACE_THREADS = 11

    def set_option(self, option: int, value: int):
        nat.check(self._lib.rflu_gpu_set_option(self.handle, int(option), int(value)),
                  "rflu_gpu_set_option")

    def get_option(self, option: int) -> int:
        v = ctypes.c_int32()
        nat.check(self._lib.rflu_gpu_get_option(self.handle, int(option), ctypes.byref(v)),
                  "rflu_gpu_get_option")
        return v.value

    def set_profiling(self, on: bool):
        nat.check(self._lib.rflu_gpu_set_profiling(self.handle, 1 if on else 0),
                  "rflu_gpu_set_profiling")

    def kernel_times(self):
        ms = (ctypes.c_double * 6)()
        cnt = (ctypes.c_int64 * 6)()
        nat.check(self._lib.rflu_gpu_get_kernel_times(self.handle, ms, cnt, 6),
                  "rflu_gpu_get_kernel_times")
        names = ["face_flux", "update", "gradient", "blend", "time_step", "bface_flux"]
        return {names[k]: (ms[k], cnt[k]) for k in range(6)}

    def close(self):
        if getattr(self, "handle", 0):
            nat.check(self._lib.rflu_gpu_destroy_region(self.handle), "rflu_gpu_destroy_region")
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuBatch(GpuRegion):
    """Independent cases (meshes with their own gas constants, CFL, dt controls and
    patches; one scheme) in one handle: every launch of the stage covers all cases,
    and each case keeps its own time, time step, end time and checks
    (``rflu_gpu_create_batch``).  State arrays are lists, one (n_cells, 5) array per
    case."""

    def __init__(self, cases: Sequence[tuple]):
        if not cases:
            raise ValueError("a batch needs at least one case")
        self.meshes = [m for m, _ in cases]
        self.nums = [n for _, n in cases]
        self.mesh, self.num = self.meshes[0], self.nums[0]
        self.sizes = [int(m.n_cells) for m in self.meshes]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.k = len(cases)
        L = nat.lib()
        descs = (nat.RegionDesc * self.k)()
        keep = [region_desc(m, num, descs[i]) for i, (m, num) in enumerate(cases)]
        h = ctypes.c_int32(0)
        nat.check(L.rflu_gpu_create_batch(self.k, descs, ctypes.byref(h)), "rflu_gpu_create_batch")
        del keep
        self.handle = h.value
        self._lib = L

    def signature(self) -> str:
        return f"batch{self.k}-N{self.n}-" + super().signature()

    def upload(self, cvs, dv=None, time: float = 0.0):
        if len(cvs) != self.k:
            raise ValueError("one state per case")
        cv = np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 5) for c in cvs])
        super().upload(cv, None, time)

    def download(self):
        cv, dv = super().download()
        o = self.offsets
        return ([cv[o[i]:o[i + 1]] for i in range(self.k)],
                [dv[o[i]:o[i + 1]] for i in range(self.k)])

    def case_stats(self, icase: int) -> dict:
        s = nat.GpuStats()
        nat.check(self._lib.rflu_gpu_get_case_stats(self.handle, int(icase), ctypes.byref(s)),
                  "rflu_gpu_get_case_stats")
        return {name: getattr(s, name) for name, _ in nat.GpuStats._fields_}

    def time_step(self, current_time=None):
        raise NotImplementedError("a batch advances through time_steps")

    def runge_kutta(self, end_time=None):
        raise NotImplementedError("a batch advances through time_steps")
