"""The oracle's threaded loops (per-cell loops and the gathers that replace the face
scatters) give the bits of its sequential loops, so the multi-core CPU baseline times
the same computation."""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import Oracle
from rocflu_amd import cases
from rocflu_amd import mesh as rm
from rocflu_amd.solver import Numerics


def run(m, num, cv0, nsteps, threads):
    orc.set_threads(threads)
    try:
        cv, dv, st = Oracle(m, num).run(cv0, nsteps)
    finally:
        orc.set_threads(1)
    return cv, dv, st


@pytest.mark.parametrize("discr,reconst,creconst", [(3, 11, 0), (4, 2, 0), (1, 2, 1)])
def test_threads_give_sequential_bits_tets(discr, reconst, creconst):
    if orc.set_threads(4) != 4:
        pytest.skip("oracle built without OpenMP")
    orc.set_threads(1)
    m = cases.tet_box(n=6, jitter=0.08, seed=4)
    num = Numerics(discr=discr, order=2, reconst=reconst, rk=4, cfl=0.5, dt_imposed=1e-3)
    num.creconst = creconst
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    a = run(m, num, cv0, 6, 1)
    b = run(m, num, cv0, 6, 4)
    assert a[0].tobytes() == b[0].tobytes()
    assert a[1].tobytes() == b[1].tobytes()
    assert a[2] == b[2]


def test_threads_give_sequential_bits_patches():
    """Outflow (subsonic) and velocity/temperature inflow patches, first- and
    second-order patch fluxes, mass in/out at stage 1."""
    if orc.set_threads(3) != 3:
        pytest.skip("oracle built without OpenMP")
    orc.set_threads(1)
    m = rm.box(rm.HEX, 12, 5, 4, jitter=0.1, seed=2,
               bc_type=(rm.BC_INFLOW_VELTEMP, rm.BC_OUTFLOW) + (rm.BC_SLIPWALL,) * 4,
               bc_opt=(rm.BCOPT_SUBSONIC, rm.BCOPT_SUBSONIC) + (0,) * 4,
               bc_pout=(0.0, 0.9) + (0.0,) * 4)
    num = Numerics(discr=3, order=2, reconst=11, rk=2, cfl=0.5, dt_imposed=1e-3)
    gp = num.gas_prime()
    rgas = gp[1]
    num.patch_vals = np.zeros((6, 5))
    num.patch_vals[0] = (0.3, 0.0, 0.0, 1.0 / rgas, 1.0)
    cv0 = cases.sod(m, num, 6.0)
    a = run(m, num, cv0, 5, 1)
    b = run(m, num, cv0, 5, 3)
    assert a[0].tobytes() == b[0].tobytes()
    assert a[2] == b[2]
