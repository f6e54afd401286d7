"""Discrete adjoint of the explicit path, for verification.

The residual R(cv, vals) of a region (FP64, first order or second order with no
gradient modifier, the Venkatakrishnan limiter or WENO-XYZ; vals are the patches' five BCDAT
values, read by inflow, farfield and injection patches) is a ``torch.autograd.Function``
whose
backward is the library's hand-written adjoint kernel (``rflu_gpu_residual_vjp``) and
whose forward-mode rule is the tangent kernel (``rflu_gpu_residual_jvp``).  SSP-RK3
steps at given time steps are composed from it in torch with the same arithmetic as
the library's update kernel, so reverse mode through ``time_integrated_wall_pressure``
is the discrete adjoint of the solver's own step, and ``torch.utils.checkpoint`` over
segments of steps gives checkpointed recomputation over long horizons.

States are device tensors [5][n_cells] in the region's internal cell order
(``DiffRegion.to_internal`` / ``to_caller`` convert).  The time steps are inputs
(recorded from a run of the solver), not functions of the state.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native as nat
from .solver import GpuRegion


class DiffRegion:
    """An FP64 region seen as a differentiable map of device tensors."""

    def __init__(self, region: GpuRegion):
        self.g = region
        self.L = region._lib
        self.h = region.handle
        self.n = region.n
        self.n_patches = int(region.mesh.n_patches)
        n2o = np.empty(self.n, dtype=np.int32)
        nat.check(self.L.rflu_gpu_cell_order(self.h, nat.ptr(n2o, ctypes.c_int32)),
                  "rflu_gpu_cell_order")
        self.new2old = torch.from_numpy(n2o.astype(np.int64))
        vol = np.empty(self.n)
        nat.check(self.L.rflu_gpu_cell_volumes(self.h, nat.ptr(vol, ctypes.c_double)),
                  "rflu_gpu_cell_volumes")
        self.dev = torch.device("cuda")
        self.vol = torch.from_numpy(vol).to(self.dev)
        self._vals = None

    # -- layout ------------------------------------------------------------------
    def to_internal(self, cv) -> torch.Tensor:
        """(n_cells, 5) caller-ordered state -> [5][n_cells] internal-order device tensor."""
        t = torch.as_tensor(np.asarray(cv, dtype=np.float64))
        return t[self.new2old].T.contiguous().to(self.dev)

    def to_caller(self, u: torch.Tensor) -> np.ndarray:
        out = np.empty((self.n, 5))
        out[self.new2old.numpy()] = u.detach().T.cpu().numpy()
        return out

    def patch_faces(self, patch: int):
        """(internal cells, areas) of one patch's boundary faces, as device tensors."""
        m = self.g.mesh
        bp = m.a["bpatch_ptr"]
        nb = int(bp[patch + 1] - bp[patch])
        cells = np.empty(nb, dtype=np.int32)
        areas = np.empty(nb)
        nat.check(self.L.rflu_gpu_patch_faces(self.h, int(patch), nat.ptr(cells, ctypes.c_int32),
                                              nat.ptr(areas, ctypes.c_double)),
                  "rflu_gpu_patch_faces")
        return (torch.from_numpy(cells.astype(np.int64)).to(self.dev),
                torch.from_numpy(areas).to(self.dev))

    # -- library calls ---------------------------------------------------------------
    def _sync(self):
        torch.cuda.current_stream().synchronize()

    def set_state(self, u: torch.Tensor):
        self._sync()
        nat.check(self.L.rflu_gpu_set_state_dev(self.h, ctypes.c_void_p(u.data_ptr())),
                  "rflu_gpu_set_state_dev")

    def set_vals(self, vals: torch.Tensor):
        v = vals.detach().to("cpu", torch.float64).contiguous().numpy().reshape(-1, 5)
        if self._vals is not None and np.array_equal(v, self._vals):
            return
        for p in range(self.n_patches):
            nat.check(self.L.rflu_gpu_set_patch_values(self.h, p, nat.ptr(np.ascontiguousarray(v[p]),
                                                                          ctypes.c_double)),
                      "rflu_gpu_set_patch_values")
        self._vals = v.copy()

    def residual_raw(self, u: torch.Tensor, vals: torch.Tensor) -> torch.Tensor:
        u = u.detach().contiguous()
        self.set_state(u)
        self.set_vals(vals)
        r = torch.empty_like(u)
        self._sync()
        nat.check(self.L.rflu_gpu_residual_dev(self.h, ctypes.c_void_p(r.data_ptr())),
                  "rflu_gpu_residual_dev")
        return r

    def jvp_raw(self, u, vals, du, dvals=None) -> torch.Tensor:
        u = u.detach().contiguous()
        self.set_state(u)
        self.set_vals(vals)
        du = du.detach().contiguous()
        dr = torch.empty_like(u)
        dv = None
        if dvals is not None:
            dv = np.ascontiguousarray(dvals.detach().to("cpu", torch.float64).numpy().reshape(-1))
        self._sync()
        nat.check(self.L.rflu_gpu_residual_jvp(self.h, ctypes.c_void_p(du.data_ptr()),
                                               None if dv is None else nat.ptr(dv, ctypes.c_double),
                                               ctypes.c_void_p(dr.data_ptr())),
                  "rflu_gpu_residual_jvp")
        return dr

    def vjp_raw(self, u, vals, mu):
        u = u.detach().contiguous()
        self.set_state(u)
        self.set_vals(vals)
        mu = mu.detach().contiguous()
        lam = torch.empty_like(u)
        dv = np.zeros(5 * self.n_patches)
        self._sync()
        nat.check(self.L.rflu_gpu_residual_vjp(self.h, ctypes.c_void_p(mu.data_ptr()),
                                               ctypes.c_void_p(lam.data_ptr()),
                                               nat.ptr(dv, ctypes.c_double)),
                  "rflu_gpu_residual_vjp")
        return lam, torch.from_numpy(dv.reshape(-1, 5))

    def residual(self, u: torch.Tensor, vals: torch.Tensor) -> torch.Tensor:
        return ResidualFn.apply(self, u, vals)


class ResidualFn(torch.autograd.Function):
    """R(cv, vals) of a DiffRegion; backward = adjoint kernel, jvp = tangent kernel."""

    @staticmethod
    def forward(dr: DiffRegion, u: torch.Tensor, vals: torch.Tensor) -> torch.Tensor:
        return dr.residual_raw(u, vals)

    @staticmethod
    def setup_context(ctx, inputs, output):
        dr, u, vals = inputs
        ctx.dr = dr
        ctx.save_for_backward(u, vals)
        ctx.save_for_forward(u, vals)

    @staticmethod
    def backward(ctx, mu):
        u, vals = ctx.saved_tensors
        lam, dvals = ctx.dr.vjp_raw(u, vals, mu)
        return None, lam, dvals.to(vals)

    @staticmethod
    def jvp(ctx, d_dr, du, dvals):
        u, vals = ctx.saved_tensors
        if du is None:
            du = torch.zeros_like(u)
        return ctx.dr.jvp_raw(u, vals, du, dvals)


def ssp_rk3_step(dr: DiffRegion, u: torch.Tensor, vals: torch.Tensor, dt: float) -> torch.Tensor:
    """One SSP-RK3 step at time step dt, in the update kernel's arithmetic
    (rk_update_cell<4>: adtv = dt/vol; q - adtv*R; 3/4 and 1/4, 1/3 and 2/3 blends)."""
    adtv = dt / dr.vol
    u1 = u - adtv * dr.residual(u, vals)
    u2 = 0.75 * u + 0.25 * (u1 - adtv * dr.residual(u1, vals))
    return (1.0 / 3.0) * u + (2.0 / 3.0) * (u2 - adtv * dr.residual(u2, vals))


def wall_pressure(dr: DiffRegion, u: torch.Tensor, cells: torch.Tensor, areas: torch.Tensor,
                  gamma_p: float) -> torch.Tensor:
    """Sum over a patch's faces of |A| p of the adjacent cell (p of RFLU_SetDependentVars)."""
    q = u[:, cells]
    ke = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0]
    p = (gamma_p - 1.0) * (q[4] - ke)
    return (areas * p).sum()


def time_integrated_wall_pressure(dr: DiffRegion, u0: torch.Tensor, vals: torch.Tensor,
                                  dts: Sequence[float], patch: int, gamma_p: float,
                                  checkpoint_every: Optional[int] = 20):
    """J = sum_n dt_n * sum_b |A_b| p(u^{n+1}) over a wall patch, after len(dts) SSP-RK3
    steps; with checkpoint_every, each segment of that many steps keeps only its input
    and is recomputed during the backward pass.  Returns (J, final state)."""
    cells, areas = dr.patch_faces(patch)

    def segment(u, vals, lo, hi):
        j = torch.zeros((), dtype=torch.float64, device=u.device)
        for n in range(lo, hi):
            u = ssp_rk3_step(dr, u, vals, dts[n])
            j = j + dts[n] * wall_pressure(dr, u, cells, areas, gamma_p)
        return u, j

    nsteps = len(dts)
    J = torch.zeros((), dtype=torch.float64, device=u0.device)
    u = u0
    step = checkpoint_every or nsteps
    for lo in range(0, nsteps, step):
        hi = min(nsteps, lo + step)
        if checkpoint_every and torch.is_grad_enabled():
            u, j = torch.utils.checkpoint.checkpoint(segment, u, vals, lo, hi, use_reentrant=False)
        else:
            u, j = segment(u, vals, lo, hi)
        J = J + j
    return J, u


def record_time_steps(region: GpuRegion, cv0, nsteps: int):
    """Runs the solver's fused loop one step at a time from cv0 and returns the time
    steps it took (global%dtMin of each step) and its final state (caller order)."""
    region.upload(cv0)
    dts = []
    for _ in range(nsteps):
        region.time_steps(1)
        dts.append(region.stats()["dt"])
    return dts, region.download()[0]
