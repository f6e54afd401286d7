"""Symmetry and periodic patches on the GPU: the one region of a grid with such patches
(rflu_mesh_build_sype) advanced by the per-callee path with borders of the region with
itself (rflu_gpu_set_borders, rflu_gpu_set_border_transform, rflu_gpu_exchange), against
the oracle's restatement of the same exchange and transforms."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rocflu_amd import cases, mesh as rm
from rocflu_amd.regions import GpuRegions, sype_part
from rocflu_amd.solver import Numerics

pytestmark = pytest.mark.gpu
L = 0.008


def run_gpu(m, num, cv0, nsteps):
    g = GpuRegions([sype_part(m)], num)
    try:
        g.upload(cv0)
        for _ in range(nsteps):
            g.runge_kutta()
        for r in g.regions:
            r.check_status()
        cv, _ = g.download()
        st = g.stats()
    finally:
        g.close()
    return cv, st


def rel(a, b):
    s = np.abs(b).max(0)
    s[1:4] = s[1:4].max()
    return np.abs(a - b).max(0) / s


@pytest.mark.parametrize("kind", ["periodic", "symmetry"])
def test_gas_jump_box(gpu, kind):
    """The reference's gas-jump box (HLLC order 2, WENO-XYZ, CRECONSTC 1, Wray RK3),
    all six patches periodic or symmetric, 100 steps."""
    g = rm.box_grid(rm.HEX, 8, 8, 8, lengths=(L,) * 3)
    if kind == "periodic":
        m = g.build_sype([110] * 6, [0] * 6, [0.0] * 6, [2, 1, 4, 3, 6, 5], [1, 1, 2, 2, 3, 3])
    else:
        m = g.build_sype([100] * 6, [0] * 6, [0.0] * 6, [0] * 6, [0] * 6)
    num = Numerics(discr=3, order=2, reconst=2, rk=2, cfl=0.4, dt_imposed=1e-9,
                   dt_min_limit=2e-10, gamma=1.4, cp=1004.64, creconst=1)
    gp = num.gas_prime()[2]
    x = m.cofg[:m.n_cells, 0]
    rho = np.where(x < 0.0043, 1.2, 0.97)
    p = np.where(x < 0.0043, 101325.0, 82700.0)
    cv0 = np.zeros((m.n_cells, 5))
    cv0[:, 0] = rho
    cv0[:, 4] = rho * (p / (rho * (gp - 1.0)) + 0.0)
    print(f"case: gas-jump {kind} box, {m.n_cells} cells + {m.n_virtual} virtual", flush=True)
    cvg, st = run_gpu(m, num, cv0, 100)
    cvo, _, sto = Oracle(m, num).run(cv0, 100)
    assert sto["rc"] == 0 and st["n_invalid"] == 0 and st["n_negative"] == 0
    err = rel(cvg, cvo)
    print("GPU vs oracle", err, flush=True)
    assert err.max() <= 1e-12


@pytest.mark.parametrize("discr,reconst,rk", [(4, 11, 4), (1, 10, 1), (3, 0, 2)])
def test_jittered_tets_with_mixed_patches(gpu, discr, reconst, rk):
    """Jittered tetrahedra, periodic in y, symmetric in z, inflow and outflow in x,
    30 steps of a Mach 2 shock with noise."""
    g = rm.box_grid(rm.TET6, 10, 4, 4, lengths=(10.0, 4.0, 4.0), jitter=0.1, seed=5)
    num = Numerics(discr=discr, order=2, reconst=reconst, rk=rk, cfl=0.5, dt_imposed=1e-3,
                   patch_order=[2, 1, 1], patch_vals=[[0.0] * 5] * 3)
    _, rgas, gp = num.gas_prime()
    m = g.build_sype([rm.BC_INFLOW_VELTEMP, rm.BC_OUTFLOW, 110, 110, 100, 100],
                     [0, 0, 0, 0, 0, 0], [0.0] * 6, [0, 0, 4, 3, 0, 0], [0, 0, 2, 2, 0, 0])
    assert list(m.a["patch_map"]) == [0, 1, -1]
    cv0 = cases.planar_shock(m, num, 3.0, mach=2.0, noise=0.01, seed=7)
    r2, u2, p2 = cases.normal_shock(2.0, gp, 1.0, 1.0)
    num.patch_vals = [[u2, 0.0, 0.0, p2 / (rgas * r2), p2], [0.0] * 5, [0.0] * 5]
    print(f"case: tets y-periodic z-symmetric discr={discr} reconst={reconst} rk={rk}", flush=True)
    cvg, st = run_gpu(m, num, cv0, 30)
    cvo, _, sto = Oracle(m, num).run(cv0, 30)
    assert sto["rc"] == 0
    err = rel(cvg, cvo)
    print("GPU vs oracle", err, flush=True)
    assert err.max() <= 1e-12
