"""The Fortran host binding (rocflu_amd/fortran/rflu_gpu.f90): a Fortran driver with
1-based, Fortran-owned region arrays calls the C-ABI in RungeKuttaMP's order and must
produce the state the Python host produces through the same entry points."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import dump as odump
from rocflu_amd import cases
from rocflu_amd.solver import GpuRegion, Numerics

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "rocflu_amd", "fortran", "rflu_driver")


@pytest.mark.parametrize("order", [1, 2])
def test_fortran_driver_matches_python_host(gpu, order):
    assert os.path.exists(DRIVER), "build the Fortran driver first (amdflang)"
    nx, ny, nz, nsteps = 32, 8, 8, 15
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "cv.bin")
        print(f"case: fortran driver order={order}", flush=True)
        p = subprocess.run([DRIVER, str(nx), str(ny), str(nz), str(nsteps), "3", str(order), out],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
        print(p.stdout, flush=True)
        assert p.returncode == 0
        cvf = np.fromfile(out, dtype=np.float64).reshape(-1, 5)
        d = odump.decode(open(out + ".dump", "rb").read())
        assert d["step"] == nsteps and d["region"] == 0 and d["cv"].tobytes() == cvf.tobytes()
    m = cases.sod_box(nx=nx, ny=ny, nz=nz, stencils=(order == 2))
    num = Numerics(discr=3, order=order, reconst=2 if order == 2 else 0, rk=2, cfl=0.5,
                   dt_imposed=1e-3)
    cv0 = cases.sod(m, num, 0.5 * nx)
    g = GpuRegion(m, num)
    g.upload(cv0)
    for _ in range(nsteps):
        g.runge_kutta()
    cvp, _ = g.download()
    g.close()
    assert cvf.shape == cvp.shape
    assert np.array_equal(cvf, cvp)
