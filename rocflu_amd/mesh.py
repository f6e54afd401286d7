"""Host-side region builder (ctypes view of ``librflu_mesh.so``).

Restates the reference's one-time region setup for grids read from CENTAUR ASCII
files (utilities/part/RFLU_ModCENTAUR.F90:776-1315) or generated as jittered boxes: the
face list (modflu/RFLU_ModFaceList.F90:784-1585), geometry (modflu/RFLU_ModGeometry.F90:
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
GRID = 2
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
        vp, i32p, f64p, i64p = (ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64))
        lib.rflu_mesh_box.restype = ctypes.c_int
        lib.rflu_mesh_box.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_uint64, i32p, i32p, f64p, ctypes.c_int,
            ctypes.POINTER(vp)]
        lib.rflu_mesh_build.argtypes = [vp, i32p, i32p, f64p, ctypes.c_int, ctypes.POINTER(vp)]
        lib.rflu_mesh_build_sype.argtypes = [vp, i32p, i32p, f64p, i32p, i32p, f64p, ctypes.c_int,
                                             ctypes.POINTER(vp)]
        lib.rflu_grid_box.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, ctypes.c_double,
            ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(vp)]
        lib.rflu_grid_read_centaur.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
        lib.rflu_grid_write_centaur.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        lib.rflu_grid_create.argtypes = [
            ctypes.c_int, f64p, ctypes.c_int64, i32p, ctypes.c_int64, i32p, ctypes.c_int64, i32p,
            ctypes.c_int64, i32p, ctypes.c_int, i32p, i32p, i32p, i32p, i32p, ctypes.POINTER(vp)]
        lib.rflu_grid_destroy.argtypes = [vp]
        lib.rflu_grid_sizes.argtypes = [vp, i64p]
        lib.rflu_grid_array.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp), i64p,
                                        ctypes.POINTER(ctypes.c_int)]
        lib.rflu_mesh_destroy.argtypes = [vp]
        lib.rflu_mesh_sizes.argtypes = [vp, i64p]
        lib.rflu_mesh_array.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp), i64p,
                                        ctypes.POINTER(ctypes.c_int)]
        lib.rflu_mesh_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {_lib().rflu_mesh_last_error().decode()}")


def _borrow(getter, h, name):
    ptr = ctypes.c_void_p()
    cnt = ctypes.c_int64()
    dt = ctypes.c_int()
    _check(getter(h, name.encode(), ctypes.byref(ptr), ctypes.byref(cnt), ctypes.byref(dt)),
           f"array {name}")
    n = cnt.value
    dtype = np.int32 if dt.value == 0 else np.float64
    if n == 0:
        return np.zeros(0, dtype=dtype)
    ctype = ctypes.c_int32 if dt.value == 0 else ctypes.c_double
    buf = ctypes.cast(ptr, ctypes.POINTER(ctype * n)).contents
    return np.frombuffer(buf, dtype=dtype).copy()


_ARRAYS = {
    "f2c": 2, "f2v": 4, "bf2c": 1, "bf2v": 4, "bpatch_ptr": 1, "patch_bc": 1,
    "patch_opt": 1, "c2v": None, "c2v_ptr": 1, "c2cs_ptr": 1, "c2cs_idx": 1, "c2cs_bptr": 1,
    "c2cs_bface": 2, "c2f_ptr": 1,
    "c2f_idx": 1, "xyz": 3, "fn": 4, "fc": 3, "bfn": 4, "bfc": 3,
    "patch_pout": 1, "vol": 1, "cofg": 3, "xyz_moms": 10,
    "virt_src": 1, "virt_xf": 1, "virt_patch": 1, "patch_map": 1, "xforms": 9,
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


def _mesh_from_handle(h, kind, dims) -> Mesh:
    """The region's arrays; virtual cells (symmetry/periodic) are the last
    len(virt_src) cells and are not counted in n_cells."""
    lib = _lib()
    try:
        sz = (ctypes.c_int64 * 9)()
        lib.rflu_mesh_sizes(h, sz)
        n_cells, n_faces, n_bfaces, n_patches, n_verts, _, nvpc = (int(v) for v in sz[:7])
        arrays = {}
        for name, ncomp in _ARRAYS.items():
            arr = _borrow(lib.rflu_mesh_array, h, name)
            if name == "c2v":
                ncomp = nvpc     # CSR (see c2v_ptr) when the cells are of several kinds
            if ncomp and ncomp > 1 and arr.size:
                arr = arr.reshape(-1, ncomp)
            arrays[name] = arr
    finally:
        lib.rflu_mesh_destroy(h)
    n_virtual = int(arrays["virt_src"].size)
    return Mesh(kind=kind, dims=dims, n_cells=n_cells - n_virtual, n_faces=n_faces,
                n_bfaces=n_bfaces, n_patches=n_patches, n_verts=n_verts, a=arrays,
                n_virtual=n_virtual)


def box(kind: int, nx: int, ny: int, nz: int, lengths=(1.0, 1.0, 1.0),
        origin=(0.0, 0.0, 0.0), jitter: float = 0.0, seed: int = 0,
        bc_type=(BC_SLIPWALL,) * 6, bc_opt=(BCOPT_SUPERSONIC,) * 6,
        bc_pout=(0.0,) * 6, stencils: bool = True) -> Mesh:
    """Build a jittered box region of hexahedra (``HEX``) or Kuhn tetrahedra (``TET6``).

    Patches are x-, x+, y-, y+, z-, z+; ``bc_type`` uses the reference codes
    (10 and 11 inflow, 20 outflow, 60 slip wall, 80 farfield, 90 injection),
    ``bc_opt`` the inflow/outflow options (0 supersonic, 1 subsonic, 2 mixed) and
    ``bc_pout`` the outflow static pressure.  The region is ``build`` of
    ``box_grid``'s grid.
    """
    lib = _lib()
    t = (ctypes.c_int32 * 6)(*[int(x) for x in bc_type])
    o = (ctypes.c_int32 * 6)(*[int(x) for x in bc_opt])
    p = (ctypes.c_double * 6)(*[float(x) for x in bc_pout])
    h = ctypes.c_void_p()
    _check(lib.rflu_mesh_box(int(kind), int(nx), int(ny), int(nz),
                             float(lengths[0]), float(lengths[1]), float(lengths[2]),
                             float(origin[0]), float(origin[1]), float(origin[2]),
                             float(jitter), ctypes.c_uint64(seed & 0xFFFFFFFFFFFFFFFF),
                             t, o, p, 1 if stencils else 0, ctypes.byref(h)), "rflu_mesh_box")
    return _mesh_from_handle(h, kind, (nx, ny, nz))


class Grid:
    """A grid as a grid file holds it (``rflu_grid``): vertices, cells per kind in the
    reference's vertex order, and per patch its triangles then quadrilaterals."""

    _NAMES = {"xyz": 3, "tet2v": 4, "hex2v": 8, "pri2v": 6, "pyr2v": 5, "patch_type": 1,
              "patch_ntri": 1, "patch_nquad": 1, "btri2v": 3, "bquad2v": 4}

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib().rflu_grid_destroy(h)
            self._h = None

    @property
    def sizes(self):
        sz = (ctypes.c_int64 * 8)()
        _check(_lib().rflu_grid_sizes(self._h, sz), "rflu_grid_sizes")
        keys = ("n_verts", "n_tets", "n_hexs", "n_pris", "n_pyrs", "n_patches", "n_btris",
                "n_bquads")
        return dict(zip(keys, (int(v) for v in sz)))

    def array(self, name: str) -> np.ndarray:
        arr = _borrow(_lib().rflu_grid_array, self._h, name)
        nc = self._NAMES[name]
        return arr.reshape(-1, nc) if nc > 1 else arr

    def write_centaur(self, path: str, field_width: int = 24) -> None:
        """Write a CENTAUR ASCII grid file (.hyb.asc)."""
        _check(_lib().rflu_grid_write_centaur(self._h, os.fsencode(path), int(field_width)),
               "rflu_grid_write_centaur")

    def build(self, bc_type, bc_opt=None, bc_pout=None, stencils: bool = True) -> Mesh:
        """The region of this grid with per-patch boundary types, options and outflow
        pressures (``rflu_mesh_build``)."""
        P = self.sizes["n_patches"]
        bc_opt = [0] * P if bc_opt is None else bc_opt
        bc_pout = [0.0] * P if bc_pout is None else bc_pout
        if not (len(bc_type) == len(bc_opt) == len(bc_pout) == P):
            raise ValueError(f"the grid has {P} patches")
        t = (ctypes.c_int32 * P)(*[int(x) for x in bc_type])
        o = (ctypes.c_int32 * P)(*[int(x) for x in bc_opt])
        p = (ctypes.c_double * P)(*[float(x) for x in bc_pout])
        h = ctypes.c_void_p()
        _check(_lib().rflu_mesh_build(self._h, t, o, p, 1 if stencils else 0, ctypes.byref(h)),
               "rflu_mesh_build")
        return _mesh_from_handle(h, GRID, None)

    def build_sype(self, bc_type, bc_opt, bc_pout, related, axis, angle=None,
                   stencils: bool = True) -> Mesh:
        """The region of this grid with symmetry (100) and periodic (110) patches as
        virtual cells (``rflu_mesh_build_sype``); ``related`` are 1-based patch numbers,
        ``axis`` 1..3, ``angle`` radians (0)."""
        P = self.sizes["n_patches"]
        angle = [0.0] * P if angle is None else angle
        args = [bc_type, bc_opt, bc_pout, related, axis, angle]
        if any(len(a) != P for a in args):
            raise ValueError(f"the grid has {P} patches")
        i32 = lambda a: (ctypes.c_int32 * P)(*[int(x) for x in a])
        f64 = lambda a: (ctypes.c_double * P)(*[float(x) for x in a])
        h = ctypes.c_void_p()
        _check(_lib().rflu_mesh_build_sype(self._h, i32(bc_type), i32(bc_opt), f64(bc_pout),
                                           i32(related), i32(axis), f64(angle),
                                           1 if stencils else 0, ctypes.byref(h)),
               "rflu_mesh_build_sype")
        return _mesh_from_handle(h, GRID, None)


def read_centaur(path: str) -> Grid:
    """Read a CENTAUR ASCII grid file (.hyb.asc) as the reference's
    RFLU_ReadGridCENTAURASCII does (utilities/part/RFLU_ModCENTAUR.F90:776-1315)."""
    h = ctypes.c_void_p()
    _check(_lib().rflu_grid_read_centaur(os.fsencode(path), ctypes.byref(h)),
           "rflu_grid_read_centaur")
    return Grid(h)


def box_grid(kind: int, nx: int, ny: int, nz: int, lengths=(1.0, 1.0, 1.0),
             origin=(0.0, 0.0, 0.0), jitter: float = 0.0, seed: int = 0) -> Grid:
    """The grid ``box`` builds its region from."""
    h = ctypes.c_void_p()
    _check(_lib().rflu_grid_box(int(kind), int(nx), int(ny), int(nz),
                                float(lengths[0]), float(lengths[1]), float(lengths[2]),
                                float(origin[0]), float(origin[1]), float(origin[2]),
                                float(jitter), ctypes.c_uint64(seed & 0xFFFFFFFFFFFFFFFF),
                                ctypes.byref(h)), "rflu_grid_box")
    return Grid(h)


def make_grid(xyz, tets=None, hexs=None, pris=None, pyrs=None, patches=()) -> Grid:
    """A grid from arrays: ``xyz`` (V, 3), cell lists (n, 4|8|6|5) 0-based in the
    reference's vertex order, and ``patches`` as a sequence of (code, triangles (n, 3),
    quadrilaterals (n, 4))."""
    def i32(a, nc):
        a = np.zeros((0, nc), np.int32) if a is None else np.asarray(a, np.int32).reshape(-1, nc)
        return np.ascontiguousarray(a)
    xyz = np.ascontiguousarray(np.asarray(xyz, np.float64).reshape(-1, 3))
    cells = [i32(tets, 4), i32(hexs, 8), i32(pris, 6), i32(pyrs, 5)]
    ptype = np.array([int(p[0]) for p in patches], np.int32)
    tris = [i32(p[1], 3) for p in patches]
    quads = [i32(p[2], 4) for p in patches]
    ntri = np.array([len(t) for t in tris], np.int32)
    nquad = np.array([len(q) for q in quads], np.int32)
    btri = np.ascontiguousarray(np.concatenate(tris) if tris else np.zeros((0, 3), np.int32))
    bquad = np.ascontiguousarray(np.concatenate(quads) if quads else np.zeros((0, 4), np.int32))
    P = ctypes.POINTER(ctypes.c_int32)

    def ptr(a):
        return a.ctypes.data_as(P)
    h = ctypes.c_void_p()
    _check(_lib().rflu_grid_create(
        len(xyz), xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        len(cells[0]), ptr(cells[0]), len(cells[1]), ptr(cells[1]),
        len(cells[2]), ptr(cells[2]), len(cells[3]), ptr(cells[3]),
        len(ptype), ptr(ptype), ptr(ntri), ptr(nquad), ptr(btri), ptr(bquad), ctypes.byref(h)),
        "rflu_grid_create")
    return Grid(h)
