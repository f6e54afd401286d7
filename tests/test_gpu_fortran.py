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


REGIONS_DRIVER = os.path.join(os.path.dirname(DRIVER), "rflu_regions_driver")


def _write_region(path, p, cv0, num):
    """One region's file in the Fortran host's layout (1-based cell numbers)."""
    m, a = p.mesh, p.mesh.a
    sends, recvs = [b.send for b in p.borders], [b.recv for b in p.borders]

    def i32(x):
        return np.ascontiguousarray(np.asarray(x), dtype=np.int32)

    def f64(x):
        return np.ascontiguousarray(np.asarray(x), dtype=np.float64)

    def cat(xs):
        return i32(np.concatenate([np.asarray(x, dtype=np.int64) for x in xs] + [np.zeros(0, np.int64)]))

    def ptr(xs):
        return i32(np.concatenate([[0], np.cumsum([len(x) for x in xs])]))

    hd = i32([m.n_cells, m.n_virtual, m.n_faces, m.n_bfaces, m.n_patches, len(a["c2cs_idx"]),
              len(p.borders), sum(len(x) for x in sends), sum(len(x) for x in recvs)])
    blocks = [hd, i32(a["f2c"]) + 1, f64(a["fn"]), f64(a["fc"]), f64(a["cofg"]), f64(a["vol"]),
              i32(a["bpatch_ptr"]), i32(a["bf2c"]) + 1, f64(a["bfn"]), f64(a["bfc"]),
              i32(a["patch_bc"]), i32(a["patch_opt"]), f64(a["patch_pout"]), i32(num.patch_order),
              i32(a["c2cs_ptr"]), i32(a["c2cs_idx"]) + 1, f64(a["xyz_moms"]), f64(num.patch_vals),
              f64(cv0[p.cells]),
              i32([b.nbr + 1 for b in p.borders]), i32([b.nbr_border + 1 for b in p.borders]),
              ptr(sends), cat(sends) + 1, ptr(recvs), cat(recvs) + 1]
    with open(path, "wb") as f:
        for x in blocks:
            f.write(x.tobytes())


def test_fortran_regions_driver_matches_python_host(gpu):
    """Three regions coupled through virtual cells, driven from Fortran (1-based
    borders, rflu_gpu_set_borders / rflu_gpu_exchange in RungeKuttaMP's places): the
    states of the Python host's loop over the same regions, bit for bit."""
    from rocflu_amd import mesh as rmesh, regions as rg
    assert os.path.exists(REGIONS_DRIVER), "build the Fortran drivers first (amdflang)"
    n, nsteps = (12, 5, 5), 10
    m = rmesh.box(rmesh.TET6, *n, lengths=tuple(float(k) for k in n), jitter=0.08, seed=5,
                  bc_type=(rmesh.BC_INFLOW_VELTEMP, rmesh.BC_OUTFLOW) + (rmesh.BC_SLIPWALL,) * 4,
                  bc_opt=(1, 1, 0, 0, 0, 0), bc_pout=(0.0, 1.0, 0, 0, 0, 0), stencils=True)
    rgas = Numerics().gas_prime()[1]
    num = Numerics(discr=3, order=2, reconst=11, rk=4, cfl=0.5, dt_imposed=1e-3,
                   patch_vals=[[0.4, 0.05, -0.02, 1.1 / rgas, 1.2]] + [[0.0] * 5] * 5,
                   patch_order=[2, 2, 1, 1, 1, 1])
    cv0 = cases.blast(m, num, (4.2, 2.5, 2.25), 1.5, p_in=4.0)
    x = m.cofg[:, 0]
    parts = rg.partition(m, np.minimum((3 * (x - x.min()) / np.ptp(x)).astype(int), 2), order=2)
    with tempfile.TemporaryDirectory() as td:
        for r, p in enumerate(parts):
            _write_region(os.path.join(td, f"in{r}.bin"), p, cv0, num)
        print(f"case: fortran regions driver, {len(parts)} regions", flush=True)
        res = subprocess.run([REGIONS_DRIVER, str(len(parts)), str(nsteps), os.path.join(td, "in"),
                              os.path.join(td, "out")], stdout=subprocess.PIPE,
                             stderr=subprocess.STDOUT, text=True, timeout=120)
        print(res.stdout, flush=True)
        assert res.returncode == 0
        cvf = np.empty((m.n_cells, 5))
        for r, p in enumerate(parts):
            cvf[p.cells[:p.n_act]] = np.fromfile(os.path.join(td, f"out{r}.bin"),
                                                 dtype=np.float64).reshape(-1, 5)
    R = rg.GpuRegions(parts, num)
    try:
        R.upload(cv0)
        for _ in range(nsteps):
            R.runge_kutta()
        cvp, _ = R.download()
    finally:
        R.close()
    assert np.array_equal(cvf, cvp)
