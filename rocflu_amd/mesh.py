"""Host-side region builder (ctypes view of ``librflu_mesh.so``).

Restates the reference's one-time region setup for jittered box meshes: the face
list (modflu/RFLU_ModFaceList.F90:784-1585), geometry (modflu/RFLU_ModGeometry.F90:
307-665), cell-to-cell stencils (modflu/RFLU_ModStencilsCells.F90:750-1416) and
R-factors (modflu/RFLU_ModWeights.F90:1770-1894).  Arrays keep the reference's
column-major (component-fastest) layout with 0-based indices.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

HEX = 0
TET6 = 1
BC_INFLOW_TOTANG = 10
BC_INFLOW_VELTEMP = 11
BC_OUTFLOW = 20
BC_SLIPWALL = 60
BC_FARFIELD = 80
BC_INJECTION = 90
BCOPT_SUPERSONIC = 0
BCOPT_SUBSONIC = 1
BCOPT_MIXED = 2


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librflu_mesh.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first")
        lib = ctypes.CDLL(path)
        lib.rflu_mesh_box.restype = ctypes.c_int
        lib.rflu_mesh_box.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_uint64,
            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
            ctypes.POINTER(ctypes.c_double), ctypes.c_int,
            ctypes.POINTER(ctypes.c_void_p)]
        lib.rflu_mesh_destroy.argtypes = [ctypes.c_void_p]
        lib.rflu_mesh_sizes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        lib.rflu_mesh_array.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int)]
        lib.rflu_mesh_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


_ARRAYS = {
    "f2c": 2, "f2v": 4, "bf2c": 1, "bf2v": 4, "bpatch_ptr": 1, "patch_bc": 1,
    "patch_opt": 1, "c2v": None, "c2cs_ptr": 1, "c2cs_idx": 1, "c2cs_bptr": 1,
    "c2cs_bface": 2, "c2f_ptr": 1,
    "c2f_idx": 1, "xyz": 3, "fn": 4, "fc": 3, "bfn": 4, "bfc": 3,
    "patch_pout": 1, "vol": 1, "cofg": 3, "xyz_moms": 10,
}


@dataclass
class Mesh:
    """A region's grid arrays (the t_grid/t_patch fields the hot path reads)."""
    kind: int
    dims: tuple
    n_cells: int
    n_faces: int
    n_bfaces: int
    n_patches: int
    n_verts: int
    a: dict = field(default_factory=dict)
    n_virtual: int = 0          # virtual cells numbered after the n_cells actual ones

    def __getattr__(self, name):
        a = self.__dict__.get("a")
        if a is not None and name in a:
            return a[name]
        raise AttributeError(name)

    @property
    def has_stencils(self) -> bool:
        return self.a["c2cs_idx"].size > 0


def box(kind: int, nx: int, ny: int, nz: int, lengths=(1.0, 1.0, 1.0),
        origin=(0.0, 0.0, 0.0), jitter: float = 0.0, seed: int = 0,
        bc_type=(BC_SLIPWALL,) * 6, bc_opt=(BCOPT_SUPERSONIC,) * 6,
        bc_pout=(0.0,) * 6, stencils: bool = True) -> Mesh:
    """Build a jittered box region of hexahedra (``HEX``) or Kuhn tetrahedra (``TET6``).

    Patches are x-, x+, y-, y+, z-, z+; ``bc_type`` uses the reference codes
    (10 and 11 inflow, 20 outflow, 60 slip wall, 80 farfield, 90 injection),
    ``bc_opt`` the inflow/outflow options (0 supersonic, 1 subsonic, 2 mixed) and
    ``bc_pout`` the outflow static pressure.
    """
    lib = _lib()
    t = (ctypes.c_int32 * 6)(*[int(x) for x in bc_type])
    o = (ctypes.c_int32 * 6)(*[int(x) for x in bc_opt])
    p = (ctypes.c_double * 6)(*[float(x) for x in bc_pout])
    h = ctypes.c_void_p()
    rc = lib.rflu_mesh_box(int(kind), int(nx), int(ny), int(nz),
                           float(lengths[0]), float(lengths[1]), float(lengths[2]),
                           float(origin[0]), float(origin[1]), float(origin[2]),
                           float(jitter), ctypes.c_uint64(seed & 0xFFFFFFFFFFFFFFFF),
                           t, o, p, 1 if stencils else 0, ctypes.byref(h))
    if rc != 0:
        raise RuntimeError(f"rflu_mesh_box failed ({rc}): "
                           f"{lib.rflu_mesh_last_error().decode()}")
    try:
        sz = (ctypes.c_int64 * 9)()
        lib.rflu_mesh_sizes(h, sz)
        n_cells, n_faces, n_bfaces, n_patches, n_verts, _, nvpc = (int(v) for v in sz[:7])
        arrays = {}
        for name, ncomp in _ARRAYS.items():
            ptr = ctypes.c_void_p()
            cnt = ctypes.c_int64()
            dt = ctypes.c_int()
            rc = lib.rflu_mesh_array(h, name.encode(), ctypes.byref(ptr), ctypes.byref(cnt),
                                     ctypes.byref(dt))
            if rc != 0:
                raise RuntimeError(lib.rflu_mesh_last_error().decode())
            n = cnt.value
            ctype = ctypes.c_int32 if dt.value == 0 else ctypes.c_double
            if n == 0:
                arr = np.zeros(0, dtype=np.int32 if dt.value == 0 else np.float64)
            else:
                buf = ctypes.cast(ptr, ctypes.POINTER(ctype * n)).contents
                arr = np.frombuffer(buf, dtype=np.int32 if dt.value == 0 else np.float64).copy()
            if name == "c2v":
                ncomp = nvpc
            if ncomp and ncomp > 1 and arr.size:
                arr = arr.reshape(-1, ncomp)
            arrays[name] = arr
    finally:
        lib.rflu_mesh_destroy(h)
    return Mesh(kind=kind, dims=(nx, ny, nz), n_cells=n_cells, n_faces=n_faces,
                n_bfaces=n_bfaces, n_patches=n_patches, n_verts=n_verts, a=arrays)
