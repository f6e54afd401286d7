"""Restatement of the reference's fluid-dump (restart file) layout, for checking.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's dump
codec; the product never imports it.

Layout (modflu/RFLU_ModRestartData.F90:7-33, writer :136-215, reader :693-872),
stream access, native (little-endian) byte order, no record markers:
  offset 0   char[8]  'RFLUDUMP'
         8   int32    version (5; 4 is still read)
        12   int64    step
        20   float64  time
        28   int32    region
        32   int32    nCells
        36   int32    nCv
        40   int32    form (1 pure, 2 volume-weighted)
        44   int32    withPicl
        48   float64  cv(iCv, 1:nCells) for iCv = 1..nCv, one variable after the other
  then, with withPicl = 1: version 4: float64 piclVF(nCells) and 3 gradient fields
  of nCells; version 5: int32 nNonzero, int32 cells(nNonzero), float64 values(nNonzero),
  int32 withGrad, 3*withGrad gradient fields of nCells;
  then int32 nPatchesNSCBC and per patch int32 global patch, nBFaces and float64
  face states cv(iCv, 1:nBFaces) one variable after the other.
"""
from __future__ import annotations

import struct

import numpy as np

_HEAD = struct.Struct("<8siqdiiiii")   # 48 bytes


def encode(cv, step, time, region, version=5, form=1):
    """Bytes of a gas-only dump of cv (n_cells, n_cv): no PICL fields, no NSCBC patches."""
    cv = np.asarray(cv, dtype="<f8")
    n_cells, n_cv = cv.shape
    head = _HEAD.pack(b"RFLUDUMP", version, step, time, region, n_cells, n_cv, form, 0)
    return head + np.ascontiguousarray(cv.T).tobytes() + struct.pack("<i", 0)


def decode(buf):
    """Parses a dump into a dict: the header fields, cv (n_cells, n_cv), the PICL fields
    (piclVF of every cell, gradient fields or None) and the NSCBC patches as a list of
    (global patch, face states (nBFaces, nCv))."""
    magic, version, step, time, region, n_cells, n_cv, form, with_picl = _HEAD.unpack_from(buf, 0)
    pos = _HEAD.size

    def f64(n):
        nonlocal pos
        a = np.frombuffer(buf, dtype="<f8", count=n, offset=pos)
        pos += 8 * n
        return a

    def i32(n):
        nonlocal pos
        a = np.frombuffer(buf, dtype="<i4", count=n, offset=pos)
        pos += 4 * n
        return a

    out = dict(magic=magic, version=version, step=step, time=time, region=region,
               n_cells=n_cells, n_cv=n_cv, form=form, with_picl=with_picl)
    out["cv"] = f64(n_cv * n_cells).reshape(n_cv, n_cells).T.copy()
    vf, grad = None, None
    if with_picl == 1:
        if version == 4:
            vf = f64(n_cells).copy()
            grad = f64(3 * n_cells).reshape(3, n_cells).copy()
        else:
            nnz = int(i32(1)[0])
            cells = i32(nnz).copy()
            vals = f64(nnz)
            vf = np.zeros(n_cells)
            vf[cells - 1] = vals
            if int(i32(1)[0]) == 1:
                grad = f64(3 * n_cells).reshape(3, n_cells).copy()
    out["picl_vf"], out["picl_grad"] = vf, grad
    patches = []
    for _ in range(int(i32(1)[0])):
        g, nbf = (int(x) for x in i32(2))
        patches.append((g, f64(n_cv * nbf).reshape(n_cv, nbf).T.copy()))
    out["nscbc"] = patches
    out["size"] = pos
    return out
