"""Regions of one problem on one device, coupled through virtual cells.

The reference runs a problem as several regions (``rflupart`` splits the grid; each
region holds its actual cells, the virtual cells mirroring its neighbours' cells, and
one border per neighbouring region, modflu/RFLU_ModCommLists.F90), and ``RungeKuttaMP``
exchanges the gradients of the cells mirrored after the gradient pass
(libfloflu/RungeKuttaMP.F90:229-236) and their state after the update (:440-459).

``partition`` builds such regions from a region grid and a cell-to-region map:

* actual cells of region r: the cells mapped to r, in grid order;
* first-layer virtual cells: every other cell that an actual cell reaches through a
  face or (order 2) through its stencil; second-layer virtual cells (order 2): every
  further cell that a first-layer cell's stencil reaches; each layer in grid order.
  First-layer cells keep their whole stencils (second-layer cells none), so the raw
  gradients they form are those of the cells they mirror -- the WENO blend of an
  actual cell reads them before the gradient exchange (without the blend, blend=False
  keeps the first layer only, without stencils);
* interior faces: the grid's faces with an actual cell, in grid order and
  orientation; boundary faces: those of actual cells, patch by patch in grid order;
* borders: one per neighbouring region s (ascending s): the actual cells that are
  virtual in s (send) and the virtual cells that s owns (receive), both in grid order,
  which makes a send list and the matching receive list agree.

The per-cell face and stencil orders are those of the grid, so a multi-region run has
the single region's arithmetic: the same states bit for bit.  (The reference builds
the virtual cells' stencils from the region's own cells, which can make its WENO
results near borders depend on the partition; that construction is a preprocessing
step, outside this path.)

``sype_part`` makes the one region of a grid with symmetry/periodic patches, whose
virtual cells mirror its own actual cells through borders of the region with itself.

``GpuRegions`` is the host loop of RungeKuttaMP over the regions, through the
per-callee entry points and ``rflu_gpu_exchange``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _native as nat
from .mesh import Mesh
from .solver import RECONST_WENO_XYZ, GpuRegion, Numerics, n_stages

XCHG_STATE = 1
XCHG_GRAD = 2


@dataclass
class Border:
    nbr: int                 # neighbouring region (index in the partition)
    nbr_border: int          # index of the matching border in the neighbour
    send: np.ndarray         # local actual cells sent to the neighbour
    recv: np.ndarray         # local virtual cells received from it


@dataclass
class RegionPart:
    mesh: Mesh               # n_cells actual cells, then mesh.n_virtual virtual cells
    cells: np.ndarray        # grid cell of every local cell (actual, then virtual)
    borders: List[Border] = field(default_factory=list)
    order: int = 2           # the order the virtual layers were built for
    blend: bool = True       # whether the virtual cells next to actual cells keep stencils
    # per border, the 3x3 turn (row-major) of what it receives, or None: a plain copy
    xforms: Optional[List[Optional[np.ndarray]]] = None

    @property
    def n_act(self) -> int:
        return int(self.mesh.n_cells)


def _csr_rows(ptr: np.ndarray, idx: np.ndarray, rows: np.ndarray):
    """The CSR rows `rows` of (ptr, idx) as a new CSR (ptr2, idx2)."""
    ptr = np.asarray(ptr, dtype=np.int64)
    lens = ptr[rows + 1] - ptr[rows]
    ptr2 = np.zeros(len(rows) + 1, dtype=np.int64)
    np.cumsum(lens, out=ptr2[1:])
    total = int(ptr2[-1])
    if total == 0:
        return ptr2, np.zeros(0, dtype=idx.dtype)
    pos = np.repeat(ptr[rows] - ptr2[:-1], lens) + np.arange(total)
    return ptr2, idx[pos]


def partition(mesh: Mesh, part, order: int = 2, blend: bool = True) -> List[RegionPart]:
    """Regions of `mesh` for the cell-to-region map `part` (values 0..K-1).

    blend=False (order 2 without the WENO-XYZ blend): nothing reads a virtual cell's
    gradient before the gradient exchange replaces it, so the virtual cells are the
    first layer only, without stencils."""
    part = np.asarray(part, dtype=np.int64).reshape(-1)
    N = int(mesh.n_cells)
    if part.shape != (N,) or part.min() < 0:
        raise ValueError("part needs one region index >= 0 per cell")
    K = int(part.max()) + 1
    if np.any(np.bincount(part, minlength=K) == 0):
        raise ValueError("every region needs cells")
    if order == 2 and not mesh.has_stencils:
        raise ValueError("order 2 needs a mesh built with stencils")
    f2c = np.asarray(mesh.f2c, dtype=np.int64).reshape(-1, 2)
    sp = np.asarray(mesh.c2cs_ptr, dtype=np.int64) if order == 2 else None
    si = np.asarray(mesh.c2cs_idx, dtype=np.int64) if order == 2 else None
    bp = np.asarray(mesh.bpatch_ptr, dtype=np.int64)
    bf2c = np.asarray(mesh.bf2c, dtype=np.int64)
    fn = np.asarray(mesh.fn).reshape(-1, 4)
    fc = np.asarray(mesh.fc).reshape(-1, 3)
    bfn = np.asarray(mesh.bfn).reshape(-1, 4)
    bfc = np.asarray(mesh.bfc).reshape(-1, 3)
    cofg = np.asarray(mesh.cofg).reshape(-1, 3)
    vol = np.asarray(mesh.vol).reshape(-1)
    moms = np.asarray(mesh.xyz_moms).reshape(-1, 10)
    if moms.shape[0] != N:                     # a grid built without stencils
        moms = np.zeros((N, 10))
    P = int(mesh.n_patches)

    parts = []
    for r in range(K):
        isact = part == r
        act = np.flatnonzero(isact)
        fmask = isact[f2c[:, 0]] | isact[f2c[:, 1]]
        near = np.zeros(N, dtype=bool)
        near[f2c[fmask].ravel()] = True
        if order == 2:
            near[_csr_rows(sp, si, act)[1]] = True
        near &= ~isact
        v1 = np.flatnonzero(near)
        if order == 2 and blend:
            far = np.zeros(N, dtype=bool)
            far[_csr_rows(sp, si, v1)[1]] = True
            far &= ~isact & ~near
            v2 = np.flatnonzero(far)
        else:
            v2 = np.zeros(0, dtype=np.int64)
        cells = np.concatenate([act, v1, v2])
        g2l = np.full(N, -1, dtype=np.int64)
        g2l[cells] = np.arange(len(cells))
        fl = np.flatnonzero(fmask)
        a = {
            "f2c": g2l[f2c[fl]].astype(np.int32),
            "fn": fn[fl].copy(), "fc": fc[fl].copy(),
            "cofg": cofg[cells].copy(), "vol": vol[cells].copy(),
            "xyz_moms": moms[cells].copy(),
            "patch_bc": np.asarray(mesh.patch_bc, dtype=np.int32).copy(),
            "patch_opt": np.asarray(mesh.patch_opt, dtype=np.int32).copy(),
            "patch_pout": np.asarray(mesh.patch_pout, dtype=np.float64).copy(),
        }
        bsel, bptr = [], [0]
        for p in range(P):
            b = np.arange(bp[p], bp[p + 1])
            b = b[isact[bf2c[b]]]
            bsel.append(b)
            bptr.append(bptr[-1] + len(b))
        bsel = np.concatenate(bsel) if bsel else np.zeros(0, dtype=np.int64)
        a["bpatch_ptr"] = np.asarray(bptr, dtype=np.int32)
        a["bf2c"] = g2l[bf2c[bsel]].astype(np.int32)
        a["bfn"] = bfn[bsel].copy()
        a["bfc"] = bfc[bsel].copy()
        n_loc = len(cells)
        if order == 2:
            # actual and (with the blend) first-layer cells keep their stencils, the
            # other virtual cells none
            rows = np.concatenate([act, v1]) if blend else act
            p2, i2 = _csr_rows(sp, si, rows)
            ptr = np.concatenate([p2, np.full(n_loc - len(rows), p2[-1], dtype=np.int64)])
            a["c2cs_ptr"] = ptr.astype(np.int32)
            a["c2cs_idx"] = g2l[i2].astype(np.int32)
            assert np.all(a["c2cs_idx"] >= 0)
        else:
            a["c2cs_ptr"] = np.zeros(n_loc + 1, dtype=np.int32)
            a["c2cs_idx"] = np.zeros(0, dtype=np.int32)
        # no wall-face stencil members: constrained gradients are not taken here
        a["c2cs_bptr"] = np.zeros(n_loc + 1, dtype=np.int32)
        a["c2cs_bface"] = np.zeros((0, 2), dtype=np.int32)
        # cell -> face lists in the gather order (interior faces ascending, then
        # boundary faces), for the host side's use; the library builds its own
        F = len(fl)
        lf2c = a["f2c"].astype(np.int64)
        lbf = a["bf2c"].astype(np.int64)
        owner = np.concatenate([lf2c[:, 0], lf2c[:, 1], lbf])
        sid = np.concatenate([np.arange(1, F + 1), -np.arange(1, F + 1),
                              F + 1 + np.arange(len(lbf))])
        key = np.concatenate([np.arange(F), np.arange(F), F + np.arange(len(lbf))])
        o = np.lexsort((key, owner))
        cnt = np.bincount(owner, minlength=n_loc)
        a["c2f_ptr"] = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
        a["c2f_idx"] = sid[o].astype(np.int32)
        m = Mesh(kind=mesh.kind, dims=mesh.dims, n_cells=len(act), n_faces=F,
                 n_bfaces=int(len(bsel)), n_patches=P, n_verts=0, a=a,
                 n_virtual=int(len(v1) + len(v2)))
        parts.append(RegionPart(mesh=m, cells=cells, order=order, blend=bool(blend)))

    # borders
    owner_of = part
    virt = [rp.cells[rp.n_act:] for rp in parts]
    for r, rp in enumerate(parts):
        g2l = {int(c): i for i, c in enumerate(rp.cells)}
        nbrs = set(int(s) for s in np.unique(owner_of[virt[r]]))
        for s in range(K):
            if s != r and np.any(owner_of[virt[s]] == r):
                nbrs.add(s)
        for s in sorted(nbrs):
            send = np.sort(virt[s][owner_of[virt[s]] == r])
            recv = np.sort(virt[r][owner_of[virt[r]] == s])
            rp.borders.append(Border(nbr=s, nbr_border=-1,
                                     send=np.array([g2l[int(c)] for c in send], dtype=np.int32),
                                     recv=np.array([g2l[int(c)] for c in recv], dtype=np.int32)))
    for r, rp in enumerate(parts):
        for b in rp.borders:
            b.nbr_border = [x.nbr for x in parts[b.nbr].borders].index(r)
    return parts


def sype_part(mesh: Mesh) -> RegionPart:
    """The one region of a grid with symmetry/periodic patches (``Grid.build_sype``): its
    virtual cells take the state and final gradients of the actual cells they mirror
    through borders of the region with itself, one per transform (the first, the
    identity, a plain copy), as the reference's serial region does through its own
    comm lists and RFLU_RELP_TransformWrapper."""
    n = int(mesh.n_cells)
    src = np.asarray(mesh.a["virt_src"], dtype=np.int32)
    xf = np.asarray(mesh.a["virt_xf"], dtype=np.int32)
    mats = np.asarray(mesh.a["xforms"], dtype=np.float64).reshape(-1, 9)
    borders, xforms = [], []
    for k in np.unique(xf):
        sel = np.nonzero(xf == k)[0]
        borders.append(Border(nbr=0, nbr_border=len(borders), send=src[sel],
                              recv=(n + sel).astype(np.int32)))
        xforms.append(None if k == 0 else mats[k].copy())
    cells = np.concatenate([np.arange(n), src]).astype(np.int64)
    return RegionPart(mesh=mesh, cells=cells, borders=borders, order=2, blend=True,
                      xforms=xforms)


def check_parts(parts: List[RegionPart], num: Numerics) -> None:
    """Rejects regions whose virtual layers cannot carry `num`'s scheme: constrained
    gradients in partitioned regions (their wall-face lists end at the region); at order 2 with
    the WENO-XYZ blend an actual cell blends the raw gradients of the virtual cells in
    its stencil before the gradient exchange replaces them, so those virtual cells need
    stencils of their own (partition(..., blend=True)); a run without them would read
    gradients that were never formed."""
    for r, p in enumerate(parts):
        if num.order == 2 and num.creconst and p.xforms is None and p.mesh.n_virtual:
            raise ValueError(f"region {r}: the partitioned regions' wall-face lists stop at "
                             "the region; constrained gradients (CRECONSTC) need one region")
        if num.order > p.order:
            raise ValueError(f"region {r} was partitioned for order {p.order}, the scheme is "
                             f"order {num.order}")
        if num.order != 2 or num.reconst != RECONST_WENO_XYZ:
            continue
        m, na = p.mesh, p.n_act
        ptr = np.asarray(m.c2cs_ptr, dtype=np.int64)
        idx = np.asarray(m.c2cs_idx, dtype=np.int64)
        read = np.unique(idx[:ptr[na]])
        read = read[read >= na]
        empty = read[ptr[read + 1] == ptr[read]]
        if empty.size:
            raise ValueError(f"region {r}: {empty.size} virtual cells in actual cells' stencils "
                             "have no stencil of their own; WENO-XYZ needs "
                             "partition(..., blend=True)")


class GpuRegions:
    """The regions of one problem on the GPU, advanced as RungeKuttaMP advances its
    regions (libfloflu/RungeKuttaMP.F90:148-508, rocflu/RFLU_TimeStepping.F90:429-640)."""

    def __init__(self, parts: List[RegionPart], num: Numerics):
        check_parts(parts, num)
        self.parts = parts
        self.num = num
        self.n = int(sum(p.n_act for p in parts))
        self.regions: List[GpuRegion] = []
        try:
            for p in parts:
                self.regions.append(GpuRegion(p.mesh, num))
            self._lib = nat.lib()
            hs = [g.handle for g in self.regions]
            self._handles = np.asarray(hs, dtype=np.int32)
            for p, g in zip(parts, self.regions):
                nb = len(p.borders)
                nbr = np.asarray([hs[b.nbr] for b in p.borders], dtype=np.int32)
                nbrb = np.asarray([b.nbr_border for b in p.borders], dtype=np.int32)
                sptr = np.concatenate([[0], np.cumsum([len(b.send) for b in p.borders])]).astype(np.int32)
                rptr = np.concatenate([[0], np.cumsum([len(b.recv) for b in p.borders])]).astype(np.int32)
                scell = np.concatenate([b.send for b in p.borders] + [np.zeros(0, np.int32)]).astype(np.int32)
                rcell = np.concatenate([b.recv for b in p.borders] + [np.zeros(0, np.int32)]).astype(np.int32)
                i32 = ctypes.c_int32
                nat.check(self._lib.rflu_gpu_set_borders(
                    g.handle, nb, nat.ptr(nbr, i32), nat.ptr(nbrb, i32), nat.ptr(sptr, i32),
                    nat.ptr(scell, i32), nat.ptr(rptr, i32), nat.ptr(rcell, i32)),
                    "rflu_gpu_set_borders")
                for b, m in enumerate(p.xforms or []):
                    if m is not None:
                        mm = np.ascontiguousarray(m, dtype=np.float64).reshape(9)
                        nat.check(self._lib.rflu_gpu_set_border_transform(
                            g.handle, b, nat.ptr(mm, ctypes.c_double)),
                            "rflu_gpu_set_border_transform")
        except Exception:
            self.close()
            raise
        self.time = 0.0

    def close(self):
        for g in self.regions:
            g.close()
        self.regions = []

    def upload(self, cv: np.ndarray, time: float = 0.0):
        """cv of every grid cell; each region takes its actual and virtual cells."""
        cv = np.asarray(cv, dtype=np.float64).reshape(-1, 5)
        for p, g in zip(self.parts, self.regions):
            g.upload(cv[p.cells], time=time)
        self.time = float(time)
        if any(p.borders for p in self.parts):
            # the virtual cells as the exchange sets them (turned copies included)
            self.exchange(XCHG_STATE)

    def download(self):
        """(cv, dv) of every grid cell from the regions' actual cells."""
        cv, dv = np.empty((self.n, 5)), np.empty((self.n, 3))
        for p, g in zip(self.parts, self.regions):
            c, d = g.download()
            cv[p.cells[:p.n_act]] = c[:p.n_act]
            dv[p.cells[:p.n_act]] = d[:p.n_act]
        return cv, dv

    def exchange(self, what: int):
        nat.check(self._lib.rflu_gpu_exchange(len(self.regions),
                                              nat.ptr(self._handles, ctypes.c_int32), int(what)),
                  "rflu_gpu_exchange")

    def time_step(self) -> float:
        """The minimum over the regions (RFLU_MinimumTimeStep's MPI_MIN all-reduce)."""
        return min(g.time_step(self.time) for g in self.regions)

    def runge_kutta(self, end_time: Optional[float] = None):
        dt = self.time_step()
        finished = False
        if end_time is not None and self.time + dt > end_time:
            dt = end_time - self.time
            finished = True
        for s in range(1, n_stages(self.num.rk) + 1):
            for g in self.regions:
                g.rk_init(s)
                g.cell_gradients()
            if self.num.order == 2:
                self.exchange(XCHG_GRAD)
            # every region's work is queued on its own stream before the first check
            # waits for one of them, so the regions run side by side
            for g in self.regions:
                g.convective_fluxes(s)
                g.rk_update(s, dt)
                g.set_dependent_vars()
            for g in self.regions:
                g.after_update(s)
            self.exchange(XCHG_STATE)
        self.time = self.time + dt
        for g in self.regions:
            g.time = self.time
        return dt, finished

    def stats(self) -> dict:
        """Counters summed over the regions (mass in/out of the stage-1 passes)."""
        out = {}
        for g in self.regions:
            for k, v in g.stats().items():
                if k in ("time", "dt", "steps", "finished", "pad"):
                    out[k] = v
                else:
                    out[k] = out.get(k, 0) + v
        return out
