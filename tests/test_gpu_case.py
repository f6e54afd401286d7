"""A case run from the reference's own input files: the template rectshktb.inp
(HLLC order 2, WENO-XYZ, constrained wall gradients, Wray RK3, subsonic
velocity/temperature inflow, the shock-tube initial state of INITFLOW FLAG 3) with its
.bc file and a CENTAUR ASCII grid written by rflu_grid_write_centaur.  The compiled
Fortran host (rocflu_amd/fortran/rflu_case_driver) reads the three files and drives
the per-callee GPU path; the Python host reads them through rocflu_amd.inputs and
drives the same entry points: both give the same bits, and the oracle the same state
to 1e-12.  The template's y and z patches are periodic (virtual cells exchanged through
borders of the region with itself); a variant makes them slip walls."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import dump as odump
from oracle.oracle import Oracle
from rocflu_amd import inputs as ri
from rocflu_amd import mesh as rm

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TEMPLATES = os.path.join(HERE, "golden", "templates")
DRIVER = os.path.join(os.path.dirname(HERE), "rocflu_amd", "fortran", "rflu_case_driver")


def slip_walls_for_periodic(bc_text):
    """The template's BC_PERIODIC sections as BC_SLIPW sections of the same patches."""
    out = []
    for sec in re.split(r"(?m)^(?=# )", bc_text):
        if sec.startswith("# BC_PERIODIC"):
            patch = re.search(r"(?m)^PATCH.*$", sec).group(0)
            sec = f"# BC_SLIPW\n{patch}\n#\n\n"
        out.append(sec)
    return "".join(out)


def make_case(tmp, kind, periodic=True):
    with open(os.path.join(TEMPLATES, "rectshktb.inp")) as f:
        inp = f.read()
    with open(os.path.join(TEMPLATES, "rectshktb.bc")) as f:
        bc = f.read() if periodic else slip_walls_for_periodic(f.read())
    with open(os.path.join(tmp, "rectshktb.inp"), "w") as f:
        f.write(inp)
    with open(os.path.join(tmp, "rectshktb.bc"), "w") as f:
        f.write(bc)
    n = (24, 3, 3) if kind == rm.HEX else (12, 2, 2)
    g = rm.box_grid(kind, *n, lengths=(0.048, 0.006, 0.006), origin=(-0.024, 0.0, 0.0),
                    jitter=0.1, seed=4)
    g.write_centaur(os.path.join(tmp, "rectshktb.hyb.asc"))


@pytest.mark.parametrize("kind,periodic", [(rm.HEX, True), (rm.TET6, True), (rm.HEX, False),
                                           (rm.TET6, False)])
def test_case_from_input_files(gpu, tmp_path, kind, periodic):
    assert os.path.exists(DRIVER), "build the Fortran drivers first (amdflang)"
    make_case(str(tmp_path), kind, periodic)
    nsteps = 20
    out = os.path.join(tmp_path, "cv.bin")
    print(f"case: rectshktb from input files, kind={kind}, periodic={periodic}, {nsteps} steps",
          flush=True)
    p = subprocess.run([DRIVER, str(tmp_path), "rectshktb", str(nsteps), out],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(p.stdout, flush=True)
    assert p.returncode == 0
    cvf = np.fromfile(out, dtype=np.float64).reshape(-1, 5)
    d = odump.decode(open(out + ".dump", "rb").read())
    assert d["step"] == nsteps and d["cv"].tobytes() == cvf.tobytes()

    case = ri.load_case("rectshktb", str(tmp_path))
    assert case.num.creconst == 1 and case.num.patch_order[0] == 2
    side = 110 if periodic else 60
    assert [b.bc_type for b in case.bcs] == [11, 20] + [side] * 4
    assert (case.mesh.n_virtual > 0) == periodic
    cv0 = ri.initial_state(case, "rectshktb")
    assert (cv0[:, 1] > 0).any() and (cv0[:, 1] == 0).any()     # the shock is in the box
    g = ri.gpu_host(case)
    try:
        g.upload(cv0)
        for _ in range(nsteps):
            g.runge_kutta(end_time=case.inp.max_time)
        for r in getattr(g, "regions", [g]):
            r.check_status()
        cvp, _ = g.download()
        time = g.time
    finally:
        g.close()
    assert np.array_equal(cvf, cvp)
    assert d["time"] == time

    cvo, _, st = Oracle(case.mesh, case.num).run(cv0, nsteps)
    assert st["rc"] == 0
    err = [np.linalg.norm(cvp[:, k] - cvo[:, k]) / np.linalg.norm(cvo[:, k]) for k in (0, 1, 4)]
    print("GPU vs oracle rel L2 (rho, rho u, rho E):", err, flush=True)
    assert max(err) <= 1e-12
