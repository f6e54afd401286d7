"""The reference's user input files on the host: ``<case>.inp``, ``<case>.bc`` and the
CENTAUR grid ``<case>.hyb.asc`` (ctypes view of ``librflu_mesh.so``, include/rflu_input.h).

``load_case`` turns the three into a region and its ``Numerics`` the way the
reference's set-up does: the .inp's NUMERICS / TIMESTEP / REFERENCE sections
(libfloflu/ReadNumericsSection.F90, ReadTimestepSection.F90, ReadReferenceSection.F90),
the .bc file's patch sections (modflu/RFLU_ModReadBcInputFile.F90) applied to the grid's
patches by number, and the grid (utilities/part/RFLU_ModCENTAUR.F90:776-1315).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import mesh as rm
from .solver import Numerics

I32, F64 = ctypes.c_int32, ctypes.c_double

BCT_INFLOW_TOTANG, BCT_INFLOW_VELTEMP, BCT_OUTFLOW, BCT_SLIPWALL = 10, 11, 20, 60
BCT_NOSLIPWALL_HFLUX, BCT_NOSLIPWALL_TEMP, BCT_FARFIELD, BCT_INJECTION = 70, 71, 80, 90
BCT_SYMMETRY, BCT_PERIODIC, BCT_VIRTUAL = 100, 110, 130


class Inp(ctypes.Structure):
    """rflu_inp: the .inp values of the path, with the reference's defaults."""
    _fields_ = [
        ("grid_format", I32), ("solution_format", I32), ("grid_source", I32),
        ("flow_model", I32), ("move_grid", I32), ("gas_model", I32),
        ("cfl", F64), ("discr", I32), ("order", I32), ("eps_entr", F64), ("diss_fact", F64),
        ("reconst", I32), ("dimens", I32), ("creconst_cells", I32), ("creconst_faces", I32),
        ("creconst_cells_weight", F64), ("creconst_faces_weight", F64), ("toler_ict", F64),
        ("stencil_dimens_cells", I32), ("stencil_dimens_faces", I32),
        ("stencil_dimens_bfaces", I32), ("order_bfaces", I32), ("axi_flag", I32),
        ("sigma", F64), ("kp", F64), ("ku", F64), ("minf", F64),
        ("flow_type", I32), ("solver_type", I32), ("rk_scheme", I32), ("start_step", I32),
        ("max_iter", I32), ("write_iter", I32), ("print_iter", I32),
        ("dt_imposed", F64), ("max_time", F64), ("write_time", F64), ("print_time", F64),
        ("dt_min_limit", F64), ("res_tol", F64),
        ("ref_velocity", F64), ("ref_pressure", F64), ("ref_density", F64), ("ref_cp", F64),
        ("ref_gamma", F64), ("ref_length", F64),
        ("init_flag", I32), ("init_vel", F64 * 3), ("init_press", F64), ("init_dens", F64),
        ("init_temp", F64), ("init_ival", I32 * 19), ("init_rval", F64 * 25),
        ("sections", ctypes.c_uint32),
    ]


class BcPatch(ctypes.Structure):
    """rflu_bc_patch: one patch's boundary condition from the .bc file."""
    _fields_ = [
        ("bc_type", I32), ("opt", I32), ("opt2", I32), ("order", I32), ("creconst", I32),
        ("related", I32), ("axis", I32), ("angle", F64), ("vals", F64 * 5), ("pout", F64),
        ("name", ctypes.c_char * 80),
    ]


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        lib = rm._lib()
        lib.rflu_read_inp.argtypes = [ctypes.c_char_p, ctypes.POINTER(Inp)]
        lib.rflu_read_bc.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(BcPatch)]
        lib.rflu_input_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


class InputError(RuntimeError):
    pass


def read_inp(path: str) -> Inp:
    """Read a .inp file (ReadInputFile's sections of the path; defaults elsewhere)."""
    out = Inp()
    rc = _lib().rflu_read_inp(os.fsencode(path), ctypes.byref(out))
    if rc != 0:
        raise InputError(f"rflu_read_inp ({rc}): {_lib().rflu_input_last_error().decode()}")
    return out


def read_bc(path: str, n_patches: int, order_bfaces: int = 1) -> list:
    """Read a .bc file for a grid of ``n_patches`` patches (1-based patch numbers)."""
    out = (BcPatch * n_patches)()
    rc = _lib().rflu_read_bc(os.fsencode(path), int(n_patches), int(order_bfaces), out)
    if rc != 0:
        raise InputError(f"rflu_read_bc ({rc}): {_lib().rflu_input_last_error().decode()}")
    return list(out)


# RKSCHEME codes of the reference (modfloflu/ModParameters.F90:615-618) are the solver's
RK_FROM_INP = {1: 1, 2: 2, 3: 3}


def numerics(inp: Inp, bcs) -> Numerics:
    """The path's Numerics from .inp and .bc values."""
    if inp.rk_scheme not in RK_FROM_INP:
        raise InputError(f"RKSCHEME {inp.rk_scheme} is not a scheme of the explicit path")
    return Numerics(
        discr=int(inp.discr), order=int(inp.order), reconst=int(inp.reconst),
        rk=RK_FROM_INP[inp.rk_scheme], cfl=float(inp.cfl), dt_imposed=float(inp.dt_imposed),
        dt_min_limit=float(inp.dt_min_limit), eps_entr=float(inp.eps_entr),
        diss_fact=float(inp.diss_fact), gamma=float(inp.ref_gamma), cp=float(inp.ref_cp),
        patch_order=[int(b.order) for b in bcs],
        patch_vals=[[float(v) for v in b.vals] for b in bcs],
        patch_opt2=[int(b.opt2) for b in bcs],
        creconst=int(inp.creconst_cells), creconst_weight=float(inp.creconst_cells_weight))


@dataclass
class Case:
    """A case as the reference's input files describe it."""
    mesh: rm.Mesh
    num: Numerics
    inp: Inp
    bcs: list
    grid: rm.Grid


def load_case(casename: str, directory: str = ".", grid_path: str | None = None,
              stencils: bool | None = None) -> Case:
    """Read ``<casename>.inp``, the grid (``<casename>.hyb.asc`` unless ``grid_path``)
    and ``<casename>.bc`` from ``directory`` and build the region."""
    base = os.path.join(directory, casename)
    inp = read_inp(base + ".inp")
    if inp.grid_source != 0:
        raise InputError(f"GRIDSRC {inp.grid_source}: only CENTAUR ASCII grids (0) are read")
    grid = rm.read_centaur(grid_path or base + ".hyb.asc")
    n_patches = grid.sizes["n_patches"]
    bcs = read_bc(base + ".bc", n_patches, inp.order_bfaces)
    for i, b in enumerate(bcs):
        if b.bc_type == 0:
            raise InputError(f"patch {i + 1} has no section in {base}.bc")
    if stencils is None:
        stencils = inp.order == 2
    types = [b.bc_type for b in bcs]
    if any(t in (BCT_SYMMETRY, BCT_PERIODIC) for t in types):
        # symmetry/periodic patches become virtual cells; the region's patches are the
        # kept ones (patch_map) and the virtual layers' outer faces (-1)
        mesh = grid.build_sype(types, [b.opt for b in bcs], [b.pout for b in bcs],
                               [b.related for b in bcs], [b.axis for b in bcs],
                               [b.angle for b in bcs], stencils=stencils)
        kept = [bcs[p] if p >= 0 else None for p in mesh.a["patch_map"]]
        num = numerics(inp, bcs)
        num.patch_order = [b.order if b else 1 for b in kept]
        num.patch_vals = [[float(v) for v in b.vals] if b else [0.0] * 5 for b in kept]
        num.patch_opt2 = [int(b.opt2) if b else 0 for b in kept]
    else:
        mesh = grid.build(types, [b.opt for b in bcs], [b.pout for b in bcs], stencils=stencils)
        num = numerics(inp, bcs)
    return Case(mesh=mesh, num=num, inp=inp, bcs=bcs, grid=grid)


def gpu_host(case: Case):
    """The GPU host of a case: a GpuRegion, or for a region with symmetry/periodic
    virtual cells the one-region GpuRegions that exchanges them."""
    from .regions import GpuRegions, sype_part
    from .solver import GpuRegion
    if case.mesh.n_virtual:
        return GpuRegions([sype_part(case.mesh)], case.num)
    return GpuRegion(case.mesh, case.num)


INITFLOW_FROMSCRATCH, INITFLOW_FROMHARDCODE = 1, 3


def initial_state(case: Case, casename: str):
    """The initial state the reference's rfluinit writes for this case
    (utilities/init): INITFLOW FLAG 1, a uniform state from VELX/VELY/VELZ/PRESS and
    DENS (or TEMP, with rho = p/(R T)); FLAG 3 for the rectangular shock tube
    ``rectshktb`` (utilities/init/RFLU_InitFlowHardCode.F90:839-901): the shocked
    state (RVAL2 density, RVAL3 x velocity, RVAL4 pressure) where x < RVAL1, the
    quiescent one (RVAL5, RVAL6) elsewhere, density times 1 + RVAL7 x + RVAL8 y +
    RVAL9 z when IVAL1 = 1.  Energies are d * MixtPerf_Eo_DGPUVW with the mixture's
    gamma (libfloflu/MixtPerf_E.F90:70-81, libflu/RFLU_SetGasVars.F90:172-178)."""
    inp, m = case.inp, case.mesh
    _, rgas, g = case.num.gas_prime()
    n = m.n_cells                      # the actual cells (virtual ones follow them)
    x, y, z = m.cofg[:n, 0], m.cofg[:n, 1], m.cofg[:n, 2]
    if inp.init_flag == INITFLOW_FROMSCRATCH:
        p = np.full(n, inp.init_press)
        u, v, w = (np.full(n, inp.init_vel[k]) for k in range(3))
        d = np.full(n, inp.init_dens) if inp.init_dens > 0 else p / (rgas * inp.init_temp)
    elif inp.init_flag == INITFLOW_FROMHARDCODE and casename == "rectshktb":
        r = inp.init_rval
        shocked = x < r[0]
        d = np.where(shocked, r[1], r[4])
        u = np.where(shocked, r[2], 0.0)
        v = np.zeros(n)
        w = np.zeros(n)
        p = np.where(shocked, r[3], r[5])
        if inp.init_ival[0] == 1:
            d = d * (1.0 + r[6] * x + r[7] * y + r[8] * z)
    else:
        raise InputError(f"INITFLOW FLAG {inp.init_flag} for case {casename!r} is not built")
    cv = np.empty((n, 5))
    cv[:, 0] = d
    cv[:, 1] = d * u
    cv[:, 2] = d * v
    cv[:, 3] = d * w
    cv[:, 4] = d * (p / (d * (g - 1.0)) + 0.5 * (u * u + v * v + w * w))
    return cv
