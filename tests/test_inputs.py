"""The .inp and .bc readers (include/rflu_input.h) against the reference's own template
case files (tests/templates/rectshktb.inp and .bc, copied under tests/golden/templates)
and against the reading rules of libfloflu/ReadInputUtil.F90 (ReadSection,
ReadPatchSection, LineHasKey) and the section readers: first occurrence wins in .inp
sections, the last in .bc sections, a key must be followed by a blank, a section ends
at (and consumes) the next line starting with '#', defaults of
libflu/RFLU_InitInputValues.F90 and RFLU_InitGlobal.F90 stay where a file is silent,
angles are converted to radians, missing switches and values are errors."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from rocflu_amd import inputs as ri

HERE = os.path.dirname(os.path.abspath(__file__))
TEMPLATES = os.path.join(HERE, "golden", "templates")
ROOT = os.path.dirname(HERE)


def write(tmp_path, name, text):
    p = os.path.join(tmp_path, name)
    with open(p, "w") as f:
        f.write(text)
    return p


def test_template_inp_values():
    inp = ri.read_inp(os.path.join(TEMPLATES, "rectshktb.inp"))
    assert inp.grid_source == 0 and inp.gas_model == 1
    assert inp.cfl == 0.4 and inp.discr == 3 and inp.order == 2 and inp.eps_entr == 0.05
    assert inp.reconst == 2 and inp.creconst_cells == 1 and inp.creconst_faces == 1
    assert inp.creconst_cells_weight == 1.0 and inp.dimens == 3
    assert (inp.sigma, inp.kp, inp.ku, inp.minf) == (1.0, 0.25, 0.75, 0.1)
    assert inp.flow_type == 1 and inp.solver_type == 1 and inp.rk_scheme == 2
    assert inp.max_time == 5.0e-4 and inp.write_time == 5.0e-6 and inp.print_time == 5.0e-6
    assert inp.dt_imposed == 1.0e-9 and inp.dt_min_limit == 2.0e-10 and inp.start_step == 0
    assert inp.ref_gamma == 1.4 and inp.ref_cp == 1004.64
    assert inp.init_flag == 3 and inp.init_ival[0] == 0
    assert list(inp.init_rval[:6]) == [-0.01, 1.4190765531700, 138.9205016483263,
                                        141869.8040079374623, 0.9700779339532, 82700.0]
    # sections not in the file keep the reference's defaults
    assert inp.ref_pressure == 1.0e5 and inp.ref_density == 1.2 and inp.toler_ict == 1e-11
    assert inp.order_bfaces == 1 and inp.max_iter == 10000 and inp.res_tol == 1e-5
    assert inp.sections == 0x7F


def test_template_bc_values():
    bcs = ri.read_bc(os.path.join(TEMPLATES, "rectshktb.bc"), 6, 1)
    b = bcs[0]
    assert b.bc_type == ri.BCT_INFLOW_VELTEMP and b.opt == 1 and b.order == 2
    # subsonic inflow: four values, no pressure (RFLU_ReadBcInflowVelTempSection nData 4)
    assert list(b.vals) == [138.9205016483263, 0.0, 0.0, 348.2905651702251, 0.0]
    assert b.name == b"Inflow"
    b = bcs[1]
    assert b.bc_type == ri.BCT_OUTFLOW and b.opt == 0 and b.order == 1 and b.pout == 0.0
    for i, (rel, axis) in enumerate([(4, 2), (3, 2), (6, 3), (5, 3)]):
        b = bcs[2 + i]
        assert b.bc_type == ri.BCT_PERIODIC and b.related == rel and b.axis == axis
        assert b.angle == 0.0


def test_defaults_without_sections(tmp_path):
    inp = ri.read_inp(write(tmp_path, "a.inp", "# NOTHING\nCFL 9\n#\n"))
    # RFLU_InitInputValues.F90:182-198, RFLU_InitGlobal.F90:229-353
    assert inp.cfl == 3.0 and inp.discr == 1 and inp.order == 1 and inp.reconst == 2
    assert inp.rk_scheme == 1 and inp.flow_type == 0 and inp.dt_imposed == 1e-5
    assert inp.max_time == 1e-4 and inp.dt_min_limit == 1e-10 and inp.print_time == 1e-12
    assert inp.ref_cp == 1004.5 and inp.ref_gamma == 1.4 and inp.sections == 0


def test_section_rules(tmp_path):
    text = "\n".join([
        "# NUMERICS ",            # header: trailing blanks do not matter
        "CFL      -0.7  ! ABS",
        "CFL      0.9",            # ReadSection: the first occurrence wins
        "ORDERX   5",              # not ORDER: the key must end at a blank
        "ORDER\t2",                # a tab ends a key too
        "DISCR 4.4",               # NINT
        "SDIMENSC 2.9",            # INT(ABS()) of an integer field
        "ENTROPY 1.0D-02",         # Fortran D exponent
        "#",
        "CFL 7",                   # outside any section: ignored
        "# TIMESTEP",
        "FLOWTYPE 1",
        "WRITIME 2.0E-3",          # PRNTIME defaults to WRITIME
        "RKSCHEME 2.6",            # INT(ABS()+0.5)
        "MAXITER 12.5",
        "# INITFLOW",              # closes TIMESTEP: this section is never read
        "FLAG 3",
        "#",
        "# INITFLOW",
        "RVAL10 5.0",              # RVAL10 is not RVAL1
        "RVAL1 -1.5",
        "IVAL1 1.9",               # an integer field truncates
        "#",
    ]) + "\n"
    inp = ri.read_inp(write(tmp_path, "b.inp", text))
    assert inp.cfl == 0.7 and inp.order == 2 and inp.discr == 4
    assert inp.stencil_dimens_cells == 2 and inp.eps_entr == 0.01
    assert inp.flow_type == 1 and inp.write_time == 2e-3 and inp.print_time == 2e-3
    assert inp.rk_scheme == 3 and inp.max_iter == 13
    assert inp.init_rval[0] == -1.5 and inp.init_rval[9] == 5.0 and inp.init_ival[0] == 1
    assert inp.init_flag == 0      # the first INITFLOW header was eaten as TIMESTEP's end


def test_unclosed_section_and_bad_value_are_errors(tmp_path):
    with pytest.raises(ri.InputError, match="closing"):
        ri.read_inp(write(tmp_path, "c.inp", "# NUMERICS\nCFL 1\n"))
    with pytest.raises(ri.InputError, match=r"c2.inp:2: no value for CFL"):
        ri.read_inp(write(tmp_path, "c2.inp", "# NUMERICS\nCFL abc\n#\n"))


BC_ALL = """! a commented section is skipped
!# BC_SLIPW
!PATCH 1 1
# BC_FARF
PATCH 1 1
CORR 0
MACH 2.0
ATTACK 30.0
SLIP  -90.0
PRESS 1.0E5
TEMP 300.0
ORDER 2
#
# BC_INFLOW
PATCH 2 2
TYPE 0
FIXED 1
PTOT 2.0E5
TTOT 350.0
BETAH 45.0
BETAV 0.0
MACH 1.5
#
# BC_OUTFLOW
PATCH 3 4
TYPE 1
PRESS 9.0E4
PRESS 8.0E4
#
# BC_INJECT
PATCH 5 5
MFRATE 2.5
TEMP 2000.0
CRECONST 1
#
# BC_SLIPW
PATCH 6 6
NAME Walls
CRECONST 1
#
# END
# BC_SLIPW
PATCH 1 6
#
"""


def test_bc_sections(tmp_path):
    bcs = ri.read_bc(write(tmp_path, "a.bc", BC_ALL), 6, 2)
    d2r = 1.0 / (180.0 / (4.0 * math.atan(1.0)))
    f = bcs[0]
    assert f.bc_type == ri.BCT_FARFIELD and f.opt2 == 0 and f.order == 2
    assert list(f.vals) == [2.0, 30.0 * d2r, -90.0 * d2r, 1e5, 300.0]
    t = bcs[1]
    assert t.bc_type == ri.BCT_INFLOW_TOTANG and t.opt == 0 and t.opt2 == 1
    assert list(t.vals) == [2e5, 350.0, 45.0 * d2r, 0.0, 1.5]
    assert t.order == 2                    # no ORDER: ORDERBF
    for o in bcs[2:4]:                     # a range of patches; the last PRESS wins
        assert o.bc_type == ri.BCT_OUTFLOW and o.opt == 1 and o.pout == 8e4
    j = bcs[4]
    assert j.bc_type == ri.BCT_INJECTION and j.creconst == 1 and list(j.vals[:2]) == [2.5, 2000.0]
    w = bcs[5]
    assert w.bc_type == ri.BCT_SLIPWALL and w.creconst == 1 and w.name == b"Walls"


@pytest.mark.parametrize("text,match", [
    ("# BC_OUTFLOW\nPATCH 1 1\nPRESS 1\n#\n", "TYPE"),
    ("# BC_OUTFLOW\nPATCH 1 1\nTYPE 1\n#\n", "PRESS"),
    ("# BC_INFLOW_VELTEMP\nPATCH 1 1\nTYPE 0\nVELX 1\nVELY 0\nVELZ 0\nTEMP 300\n#\n", "PRESS"),
    ("# BC_FARF\nPATCH 1 1\nMACH 2\n#\n", "CORR"),
    ("# BC_PERIODIC\nPATCH 1 1\nRELPATCH 2\nAXIS 1\n#\n", "ANGLE"),
    ("# BC_SLIPW\nPATCH 1 7\n#\n", "ends at 7"),
])
def test_bc_errors(tmp_path, text, match):
    with pytest.raises(ri.InputError, match=match):
        ri.read_bc(write(tmp_path, "e.bc", text), 6, 1)


def test_numerics_from_inputs():
    inp = ri.read_inp(os.path.join(TEMPLATES, "rectshktb.inp"))
    bcs = ri.read_bc(os.path.join(TEMPLATES, "rectshktb.bc"), 6, inp.order_bfaces)
    num = ri.numerics(inp, bcs)
    assert (num.discr, num.order, num.reconst, num.rk) == (3, 2, 2, 2)
    assert num.cfl == 0.4 and num.dt_imposed == 1e-9 and num.dt_min_limit == 2e-10
    assert num.gamma == 1.4 and num.cp == 1004.64 and num.creconst == 1
    assert num.patch_order == [2, 1, 1, 1, 1, 1]
    assert num.patch_vals[0] == [138.9205016483263, 0.0, 0.0, 348.2905651702251, 0.0]


def test_fortran_types_match_the_c_layout(tmp_path):
    """The BIND(C) rflu_inp / rflu_bc_patch of rocflu_amd/fortran/rflu_gpu.f90 read every
    field where the C structs put it."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not available")
    lib = os.path.join(ROOT, "rocflu_amd")
    exe = os.path.join(tmp_path, "ilc")
    subprocess.run([fc, "-O1", "-J", str(tmp_path),
                    os.path.join(lib, "fortran", "rflu_gpu.f90"),
                    os.path.join(HERE, "native", "inp_layout_check.f90"),
                    "-L" + lib, "-lrocflu_amd", "-lrflu_mesh", "-Wl,-rpath," + lib, "-o", exe],
                   check=True, cwd=str(tmp_path), capture_output=True)
    inpf, bcf = os.path.join(TEMPLATES, "rectshktb.inp"), os.path.join(TEMPLATES, "rectshktb.bc")
    out = subprocess.run([exe, inpf, bcf, "6"], check=True, capture_output=True, text=True).stdout
    vals = {}
    for line in out.splitlines():
        k, _, v = line.partition(" ")
        vals[k] = v.strip()
    inp = ri.read_inp(inpf)
    bcs = ri.read_bc(bcf, 6, inp.order_bfaces)
    for k in ("grid_source", "gas_model", "discr", "order", "reconst", "creconst_cells",
              "order_bfaces", "rk_scheme", "print_iter", "init_flag", "sections"):
        assert int(vals[k]) == getattr(inp, k), k
    for k in ("cfl", "diss_fact", "toler_ict", "minf", "dt_imposed", "max_time", "res_tol",
              "ref_cp", "ref_length", "init_temp"):
        assert float(vals[k]) == getattr(inp, k), k
    assert int(vals["init_ival19"]) == inp.init_ival[18]
    for i in range(25):
        assert float(vals[f"init_rval{i + 1}"]) == inp.init_rval[i]
    for i, b in enumerate(bcs):
        ints = [int(x) for x in vals[f"patch{i + 1}"].split()]
        assert ints == [b.bc_type, b.opt, b.opt2, b.order, b.creconst, b.related, b.axis]
        reals = [float(x) for x in vals[f"patchval{i + 1}"].split()]
        assert np.array_equal(reals, [b.angle] + list(b.vals) + [b.pout])
        name = vals.get(f"patchname{i + 1}", "").replace("\x00", " ").strip()
        assert name == b.name.decode()
