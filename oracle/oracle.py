"""ctypes view of the CPU oracle (liboracle: rflu_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker of the GPU path; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librflu_oracle.so")

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)


class OrcRegion(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int32), ("n_faces", ctypes.c_int32), ("n_patches", ctypes.c_int32),
        ("f2c", _i32p), ("fn", _f64p), ("fc", _f64p), ("cofg", _f64p), ("vol", _f64p),
        ("bpatch_ptr", _i32p), ("bf2c", _i32p), ("bfn", _f64p), ("bfc", _f64p),
        ("patch_bc", _i32p), ("patch_opt", _i32p), ("patch_pout", _f64p), ("patch_order", _i32p),
        ("c2cs_ptr", _i32p), ("c2cs_idx", _i32p), ("xyz_moms", _f64p),
        ("c2f_ptr", _i32p), ("c2f_idx", _i32p),
        ("gamma", ctypes.c_double), ("cp", ctypes.c_double),
        ("discr", ctypes.c_int32), ("order", ctypes.c_int32), ("reconst", ctypes.c_int32),
        ("rk", ctypes.c_int32),
        ("cfl", ctypes.c_double), ("dt_imposed", ctypes.c_double),
        ("dt_min_limit", ctypes.c_double), ("eps_entr", ctypes.c_double),
        ("diss_fact", ctypes.c_double),
        ("patch_vals", _f64p),
        ("c2cs_bptr", _i32p), ("c2cs_bface", _i32p),
        ("creconst_weight", ctypes.c_double), ("creconst", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("patch_opt2", _i32p),
        ("n_act", ctypes.c_int32), ("pad2_", ctypes.c_int32),
        ("virt_src", _i32p), ("virt_xf", _i32p), ("xforms", _f64p),
    ]


class OrcStats(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt_last", ctypes.c_double),
                ("steps", ctypes.c_int64), ("n_bounds_fix", ctypes.c_int64),
                ("n_pos_fix", ctypes.c_int64), ("mass_in", ctypes.c_double),
                ("mass_out", ctypes.c_double)]


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "rflu_oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.check_call(["make", "-C", HERE, "-s"])
    return LIB


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        R = ctypes.POINTER(OrcRegion)
        L.orc_run.argtypes = [R, _f64p, _f64p, ctypes.c_int64, ctypes.POINTER(OrcStats)]
        L.orc_run.restype = ctypes.c_int
        L.orc_residual.argtypes = [R, _f64p, _f64p, _f64p]
        L.orc_run_until.argtypes = [R, _f64p, _f64p, ctypes.c_double, ctypes.c_int64,
                                    ctypes.POINTER(OrcStats)]
        L.orc_run_until.restype = ctypes.c_int
        L.orc_set_dependent_vars.argtypes = [R, _f64p, _f64p, ctypes.c_int, ctypes.c_int]
        L.orc_gradients.argtypes = [R, _f64p, _f64p]
        L.orc_gradients_constr.argtypes = [R, _f64p, _f64p]
        L.orc_venkat.argtypes = [R, _f64p, _f64p]
        L.orc_weno_xyz.argtypes = [R, _f64p]
        L.orc_limit_simple.argtypes = [R, _f64p, _f64p]
        L.orc_local_dt.argtypes = [R, _f64p, _f64p]
        L.orc_local_dt.restype = ctypes.c_double
        L.orc_min_timestep.argtypes = [R, ctypes.c_double, ctypes.c_double]
        L.orc_min_timestep.restype = ctypes.c_double
        L.orc_rk_update.argtypes = [R, ctypes.c_int, ctypes.c_double, _f64p, _f64p, _f64p, _f64p]
        L.orc_after_update.argtypes = [R, _f64p, _f64p, ctypes.POINTER(OrcStats)]
        L.orc_after_update.restype = ctypes.c_int
        L.orc_fluxes.argtypes = [R, _f64p, _f64p, _f64p, _f64p, ctypes.POINTER(OrcStats),
                                 ctypes.c_int]
        L.orc_rind_state.argtypes = [R, ctypes.c_int, _f64p, _f64p, ctypes.c_double, _f64p]
        L.orc_patch_flux.argtypes = [R, ctypes.c_int, _f64p, _f64p, ctypes.c_double, _f64p]
        L.orc_interior_flux.argtypes = [R, _f64p, _f64p, _f64p, ctypes.c_int, _f64p]
        L.orc_barth_jesp.argtypes = [R, _f64p, _f64p]
        L.orc_sym_eigen3.argtypes = [_f64p, _f64p, _f64p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _L = L
    return _L


def set_threads(n: int) -> int:
    """Threads of the oracle's loops (1: the sequential restatement); returns the count
    in effect.  Any count gives the same bits."""
    L = lib()
    L.orc_set_threads(int(n))
    return L.orc_get_threads()


def _p(a, t):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


class Oracle:
    """One region for the oracle (same arrays as the GPU region)."""

    def __init__(self, mesh, num):
        from rocflu_amd.solver import region_arrays
        self.mesh, self.num = mesh, num
        self.n_act = int(mesh.n_cells)
        self.n = self.n_act + int(getattr(mesh, "n_virtual", 0))
        k = region_arrays(mesh, num)
        self._keep = k
        r = OrcRegion()
        i32, f64 = ctypes.c_int32, ctypes.c_double
        r.n_cells, r.n_faces, r.n_patches = self.n, int(mesh.n_faces), int(mesh.n_patches)
        for name in ("f2c", "bpatch_ptr", "bf2c", "patch_bc", "patch_opt", "patch_order",
                     "c2f_ptr", "c2f_idx", "patch_opt2"):
            setattr(r, name, _p(k[name], i32))
        for name in ("fn", "fc", "cofg", "vol", "bfn", "bfc", "patch_pout", "patch_vals"):
            setattr(r, name, _p(k[name], f64))
        if num.order == 2 and num.creconst:
            r.c2cs_bptr = _p(k["c2cs_bptr"], i32)
            r.c2cs_bface = _p(k["c2cs_bface"], i32)
            r.creconst, r.creconst_weight = 1, num.creconst_weight
        if num.order == 2:
            r.c2cs_ptr = _p(k["c2cs_ptr"], i32)
            r.c2cs_idx = _p(k["c2cs_idx"], i32)
            r.xyz_moms = _p(k["xyz_moms"], f64)
        r.gamma, r.cp = num.gamma, num.cp
        r.discr, r.order, r.reconst, r.rk = num.discr, num.order, num.reconst, num.rk
        r.cfl, r.dt_imposed, r.dt_min_limit = num.cfl, num.dt_imposed, num.dt_min_limit
        r.eps_entr, r.diss_fact = num.eps_entr, num.diss_fact
        if self.n > self.n_act:
            for name in ("virt_src", "virt_xf"):
                k[name] = np.ascontiguousarray(mesh.a[name], dtype=np.int32)
                setattr(r, name, _p(k[name], i32))
            k["xforms"] = np.ascontiguousarray(mesh.a["xforms"], dtype=np.float64).reshape(-1)
            r.xforms = _p(k["xforms"], f64)
            r.n_act = self.n_act
        self.r = r
        self.L = lib()

    def rind_state(self, patch, n, q, pl):
        """Boundary state of one face of `patch` (see orc_rind_state)."""
        n = np.ascontiguousarray(n, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty(6)
        self.L.orc_rind_state(ctypes.byref(self.r), int(patch), _p(n, ctypes.c_double),
                              _p(q, ctypes.c_double), float(pl), _p(out, ctypes.c_double))
        return out

    def patch_flux(self, patch, n, q, pl):
        """Flux of one face of `patch` from interior state q and pressure pl."""
        n = np.ascontiguousarray(n, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty(5)
        self.L.orc_patch_flux(ctypes.byref(self.r), int(patch), _p(n, ctypes.c_double),
                              _p(q, ctypes.c_double), float(pl), _p(out, ctypes.c_double))
        return out

    def interior_flux(self, cv, dv, grad, f):
        """Flux of interior face f alone (cv (N,5), dv (N,3), grad (N,5,3) at order 2)."""
        cv = np.ascontiguousarray(cv, dtype=np.float64)
        dv = np.ascontiguousarray(dv, dtype=np.float64)
        grad = np.ascontiguousarray(np.zeros((self.n, 5, 3)) if grad is None else grad,
                                    dtype=np.float64)
        out = np.empty(5)
        self.L.orc_interior_flux(ctypes.byref(self.r), _p(cv, ctypes.c_double),
                                 _p(dv, ctypes.c_double), _p(grad, ctypes.c_double), int(f),
                                 _p(out, ctypes.c_double))
        return out

    def dv(self, cv):
        cv = np.ascontiguousarray(cv, dtype=np.float64)
        dv = np.empty((self.n, 3))
        self.L.orc_set_dependent_vars(ctypes.byref(self.r), _p(cv, ctypes.c_double),
                                      _p(dv, ctypes.c_double), 0, self.n)
        return dv

    def _full(self, cv, dv):
        """State arrays over every cell: the actual cells' rows, then the virtual cells'
        (filled by the oracle from the cells they mirror)."""
        cv = np.asarray(cv, dtype=np.float64)
        full = np.zeros((self.n, 5))
        full[:cv.shape[0]] = cv
        if dv is None:
            d = np.zeros((self.n, 3))
            self.L.orc_set_dependent_vars(ctypes.byref(self.r), _p(full, ctypes.c_double),
                                          _p(d, ctypes.c_double), 0, self.n_act)
        else:
            d = np.zeros((self.n, 3))
            d[:np.asarray(dv).shape[0]] = dv
        return np.ascontiguousarray(full), np.ascontiguousarray(d)

    def run(self, cv, nsteps, time=0.0, dv=None):
        """nsteps of the time loop from cv (the actual cells; with virtual cells the
        result keeps the actual cells' rows)."""
        cv, dv = self._full(cv, dv)
        st = OrcStats()
        st.time = time
        rc = self.L.orc_run(ctypes.byref(self.r), _p(cv, ctypes.c_double),
                            _p(dv, ctypes.c_double), int(nsteps), ctypes.byref(st))
        stats = {name: getattr(st, name) for name, _ in OrcStats._fields_}
        stats["rc"] = rc
        return cv[:self.n_act], dv[:self.n_act], stats

    def run_until(self, cv, end_time, max_steps=10**9, time=0.0, dv=None):
        cv, dv = self._full(cv, dv)
        st = OrcStats()
        st.time = time
        rc = self.L.orc_run_until(ctypes.byref(self.r), _p(cv, ctypes.c_double),
                                  _p(dv, ctypes.c_double), float(end_time), int(max_steps),
                                  ctypes.byref(st))
        stats = {name: getattr(st, name) for name, _ in OrcStats._fields_}
        stats["rc"] = rc
        return cv[:self.n_act], dv[:self.n_act], stats

    def residual(self, cv, dv=None):
        cv = np.ascontiguousarray(cv, dtype=np.float64)
        dv = self.dv(cv) if dv is None else np.ascontiguousarray(dv, dtype=np.float64)
        rhs = np.empty((self.n, 5))
        self.L.orc_residual(ctypes.byref(self.r), _p(cv, ctypes.c_double),
                            _p(dv, ctypes.c_double), _p(rhs, ctypes.c_double))
        return rhs

    def gradients(self, var):
        var = np.ascontiguousarray(var, dtype=np.float64)
        g = np.zeros((self.n, 5, 3))
        self.L.orc_gradients(ctypes.byref(self.r), _p(var, ctypes.c_double),
                             _p(g, ctypes.c_double))
        return g

    def gradients_constr(self, var, grad):
        """Constrained refit (CRECONSTC) of the gradients of cells next to walls."""
        var = np.ascontiguousarray(var, dtype=np.float64)
        g = np.array(grad, dtype=np.float64, copy=True, order="C")
        self.L.orc_gradients_constr(ctypes.byref(self.r), _p(var, ctypes.c_double),
                                    _p(g, ctypes.c_double))
        return g

    def venkat(self, var, grad):
        var = np.ascontiguousarray(var, dtype=np.float64)
        g = np.array(grad, dtype=np.float64, copy=True, order="C")
        self.L.orc_venkat(ctypes.byref(self.r), _p(var, ctypes.c_double), _p(g, ctypes.c_double))
        return g

    def barth_jesp(self, var, grad):
        """RECONST 10: Barth-Jespersen for rho and p, the velocity gradient limited along
        the principal axes of its symmetric part; grad (N,5,3) -> limited copy."""
        var = np.ascontiguousarray(var, dtype=np.float64)
        g = np.array(grad, dtype=np.float64, copy=True, order="C")
        self.L.orc_barth_jesp(ctypes.byref(self.r), _p(var, ctypes.c_double), _p(g, ctypes.c_double))
        return g

    def weno_xyz(self, grad):
        g = np.array(grad, dtype=np.float64, copy=True, order="C")
        self.L.orc_weno_xyz(ctypes.byref(self.r), _p(g, ctypes.c_double))
        return g

    def limit_simple(self, var, grad):
        var = np.ascontiguousarray(var, dtype=np.float64)
        g = np.array(grad, dtype=np.float64, copy=True, order="C")
        self.L.orc_limit_simple(ctypes.byref(self.r), _p(var, ctypes.c_double),
                                _p(g, ctypes.c_double))
        return g

    def local_dt(self, cv, dv=None):
        cv = np.ascontiguousarray(cv, dtype=np.float64)
        dv = self.dv(cv) if dv is None else np.ascontiguousarray(dv, dtype=np.float64)
        return self.L.orc_local_dt(ctypes.byref(self.r), _p(cv, ctypes.c_double),
                                   _p(dv, ctypes.c_double))

    def min_timestep(self, dt_local_min, time):
        return self.L.orc_min_timestep(ctypes.byref(self.r), float(dt_local_min), float(time))

    def rk_update(self, stage, dt, cv, cv_old, rhs, rhs_sum):
        cv = np.ascontiguousarray(cv, dtype=np.float64)
        self.L.orc_rk_update(ctypes.byref(self.r), int(stage), float(dt), _p(cv, ctypes.c_double),
                             _p(np.ascontiguousarray(cv_old), ctypes.c_double),
                             _p(np.ascontiguousarray(rhs), ctypes.c_double),
                             _p(rhs_sum, ctypes.c_double))
        return cv

    def after_update(self, cv, dv):
        cv = np.array(cv, dtype=np.float64, copy=True, order="C")
        dv = np.array(dv, dtype=np.float64, copy=True, order="C")
        st = OrcStats()
        rc = self.L.orc_after_update(ctypes.byref(self.r), _p(cv, ctypes.c_double),
                                     _p(dv, ctypes.c_double), ctypes.byref(st))
        return rc, cv, dv, {name: getattr(st, name) for name, _ in OrcStats._fields_}


def sym_eigen3(m):
    """The limiter's Jacobi eigen solver on a symmetric 3x3 matrix: (values ascending,
    vectors as columns)."""
    a = np.asfortranarray(np.asarray(m, dtype=np.float64)).reshape(-1, order="F").copy()
    vec = np.empty(9)
    val = np.empty(3)
    lib().orc_sym_eigen3(_p(a, ctypes.c_double), _p(vec, ctypes.c_double), _p(val, ctypes.c_double))
    return val, vec.reshape(3, 3, order="F")
