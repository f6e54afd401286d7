"""Pins of the oracle's outflow patch (BC 20: bcond_outflow_perf restating
libfloflu/BcondOutflowPerf.F90:111-157, and its face flux,
rocflu/RFLU_CentralFirstPatch.F90:571-720) by what the characteristic boundary
condition fixes:

* subsonic faces impose the static pressure p_out and carry the four outgoing
  characteristics of the interior, linearised about the interior state (rho, c, n
  outward): the entropy wave dp - c^2 drho = 0, the tangential velocity du - (du.n)n = 0
  and the outgoing acoustic wave dp + rho c du.n = 0; the energy is that of p_out;
* where the linearised normal velocity would point inwards, each velocity component
  keeps the interior's sign and is at least as large as the interior's;
* supersonic faces (either option) copy the interior state;
* the face flux is the central average of the interior and boundary Euler fluxes,
  the boundary one with the pressure the option prescribes;
* on a 1-D row: an entropy wave leaves through the patch and leaves the row uniform (no
  acoustic reflection), at subsonic and supersonic speeds; an acoustic pulse meeting the
  fixed-pressure boundary returns with its pressure inverted (an open end), whereas a
  slip wall returns it with the same sign.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rocflu_amd import cases
from rocflu_amd import mesh as rm
from rocflu_amd.solver import Numerics

MM, RG, GP = Numerics().gas_prime()
POUT = 0.93


def outflow_oracle(opt, pout=POUT):
    m = rm.box(rm.HEX, 2, 2, 2, bc_type=(rm.BC_OUTFLOW,) + (rm.BC_SLIPWALL,) * 5,
               bc_opt=(opt, 0, 0, 0, 0, 0), bc_pout=(pout, 0, 0, 0, 0, 0), stencils=False)
    return Oracle(m, Numerics())


def cons(rho, u, p):
    u = np.asarray(u, dtype=float)
    return np.array([rho, *(rho * u), p / (GP - 1.0) + 0.5 * rho * u @ u])


def unit(v):
    v = np.asarray(v, dtype=float)
    return v / np.linalg.norm(v)


NORMALS = [unit([1.0, 0.0, 0.0]), unit([0.6, -0.8, 0.0]), unit([0.3, 0.5, -0.81])]


@pytest.mark.parametrize("opt", [1, 2])
@pytest.mark.parametrize("nvec", NORMALS)
@pytest.mark.parametrize("ucoef", [(0.3, 0.1, -0.05), (0.05, 0.2, 0.1)])
def test_subsonic_outflow_characteristics(opt, nvec, ucoef):
    o = outflow_oracle(opt)
    rho, p = 1.1, 1.0
    # a velocity leaving through the face (u.n > 0), Mach < 1
    u = np.asarray(ucoef) + 0.4 * nvec
    q = cons(rho, u, p)
    n4 = np.array([*nvec, 1.7])
    b = o.rind_state(0, n4, q, p)
    rb, ub, pb = b[0], b[1:4] / b[0], b[5]
    c = np.sqrt(GP * p / rho)
    dp, drho, du = pb - p, rb - rho, ub - u
    assert pb == POUT
    assert abs(b[4] - (pb / (GP - 1.0) + 0.5 * rb * ub @ ub)) <= 1e-14 * b[4]
    scale = abs(dp)
    assert abs(dp - c * c * drho) <= 1e-13 * scale              # entropy
    assert np.abs(du - (du @ nvec) * nvec).max() <= 1e-14 * scale  # tangential velocity
    assert abs(dp + rho * c * (du @ nvec)) <= 1e-13 * scale     # outgoing acoustic wave
    # the inward-normal alternative would break the acoustic relation
    assert abs(dp - rho * c * (du @ nvec)) > 0.5 * scale


def test_backflow_keeps_signs_and_magnitudes():
    """A boundary pressure above the interior's drives the linearised normal velocity
    inwards: every component keeps the interior's sign and does not shrink."""
    o = outflow_oracle(1, pout=1.3)
    rho, p = 1.0, 1.0
    nvec = NORMALS[1]
    u = np.array([0.02, 0.15, -0.07]) + 0.01 * nvec
    q = cons(rho, u, p)
    b = o.rind_state(0, np.array([*nvec, 1.0]), q, p)
    ub = b[1:4] / b[0]
    c = np.sqrt(GP * p / rho)
    lin = u + nvec * (p - 1.3) / (rho * c)
    assert lin @ nvec < 0.0
    assert np.all(np.sign(ub) == np.sign(u))
    assert np.all(np.abs(ub) >= np.abs(u)) and np.all(np.abs(ub) >= np.abs(lin) - 1e-15)
    assert b[0] == pytest.approx(rho - (p - 1.3) / (c * c), rel=1e-14)
    assert b[5] == 1.3


@pytest.mark.parametrize("opt,mach", [(1, 1.4), (2, 2.0), (0, 0.5), (0, 1.5)])
def test_supersonic_faces_copy_the_interior(opt, mach):
    o = outflow_oracle(opt)
    rho, p = 1.2, 1.0
    c = np.sqrt(GP * p / rho)
    u = mach * c * unit([1.0, 0.2, 0.0])
    q = cons(rho, u, p)
    b = o.rind_state(0, np.array([*NORMALS[0], 1.0]), q, p)
    assert np.array_equal(b[:5], q)
    # the flux pressure is the prescribed one unless the option is supersonic
    assert b[5] == (p if opt == 0 else POUT)


@pytest.mark.parametrize("opt,mach", [(1, 0.4), (1, 1.5), (0, 0.4)])
def test_outflow_flux_is_the_central_average(opt, mach):
    o = outflow_oracle(opt)
    rho, p = 0.9, 1.1
    nvec = NORMALS[2]
    u = mach * np.sqrt(GP * p / rho) * nvec + np.array([0.01, 0.02, 0.0])
    q = cons(rho, u, p)
    n4 = np.array([*nvec, 2.5])
    b = o.rind_state(0, n4, q, p)
    f = o.patch_flux(0, n4, q, p)

    def euler(s, ps):
        qn = s[1:4] @ nvec / s[0]
        return np.array([s[0] * qn, *(s[1:4] * qn + ps * nvec), (s[4] + ps) * qn])

    expect = 0.5 * (euler(q, p) + euler(b[:5], b[5])) * 2.5
    assert np.allclose(f, expect, rtol=1e-14, atol=1e-14)


def row(bc_right, opt=1, inflow=None, n=200):
    """A 1-D row of n hexes on [0, 1]; velocity/temperature inflow (or a slip wall) at
    x-, bc_right at x+ with p_out = 1."""
    left = rm.BC_INFLOW_VELTEMP if inflow is not None else rm.BC_SLIPWALL
    m = rm.box(rm.HEX, n, 1, 1, lengths=(1.0, 1.0 / n, 1.0 / n), jitter=0.0,
               bc_type=(left, bc_right) + (rm.BC_SLIPWALL,) * 4, bc_opt=(opt, opt, 0, 0, 0, 0),
               bc_pout=(0.0, 1.0, 0, 0, 0, 0), stencils=False)
    return m


@pytest.mark.parametrize("discr", [1, 3, 4])
@pytest.mark.parametrize("opt,u0", [(1, 0.5), (0, 2.0), (1, 2.0)])
def test_entropy_wave_leaves_through_the_outflow(discr, opt, u0):
    """A density blob at constant pressure in a uniform stream (Mach 0.42 or 1.7)
    crosses the outflow patch; afterwards the row is uniform and the pressure and
    velocity never moved (the patch reflects nothing)."""
    rg = Numerics(discr=discr).gas_prime()[1]
    num = Numerics(discr=discr, order=1, rk=2, cfl=0.5, dt_imposed=1e-4,
                   patch_vals=[[u0, 0.0, 0.0, 1.0 / rg, 1.0]] + [[0.0] * 5] * 5)
    m = row(rm.BC_OUTFLOW, opt, inflow=True)
    x = m.cofg[:, 0]
    rho = 1.0 + 0.1 * np.exp(-((x - 0.7) / 0.04) ** 2)
    u = np.full_like(x, u0)
    cv0 = cases.prim_to_cons(rho, u, 0 * u, 0 * u, np.ones_like(x), GP)
    cv, dv, st = Oracle(m, num).run_until(cv0, 0.6 / u0)
    assert st["rc"] == 0
    assert np.abs(cv[:, 0] - 1.0).max() < 1e-6            # the blob (0.1) has left
    assert np.abs(dv[:, 0] - 1.0).max() < 1e-12
    assert np.abs(cv[:, 1] / cv[:, 0] - u0).max() < 1e-12


@pytest.mark.parametrize("discr", [1, 3, 4])
def test_acoustic_pulse_reflects_inverted_from_fixed_pressure(discr):
    """A right-running acoustic pulse (dp = rho c du, 1e-3 of p) meets the x+ patch and
    comes back.  Fixed static pressure: reflection coefficient -1 in pressure; a slip
    wall: +1.  Compared at the same time, after the same first-order dissipation."""
    num = Numerics(discr=discr, order=1, rk=2, cfl=0.5, dt_imposed=1e-4)
    peaks = {}
    for bc in (rm.BC_OUTFLOW, rm.BC_SLIPWALL):
        m = row(bc)
        x = m.cofg[:, 0]
        c0 = np.sqrt(GP)
        dp = 1e-3 * np.exp(-((x - 0.5) / 0.05) ** 2)
        cv0 = cases.prim_to_cons(1.0 + dp / c0 ** 2, dp / c0, 0 * x, 0 * x, 1.0 + dp, GP)
        cv, dv, st = Oracle(m, num).run_until(cv0, 0.9 / c0)
        assert st["rc"] == 0
        i = int(np.argmax(np.abs(dv[:, 0] - 1.0)))
        assert abs(x[i] - 0.6) < 0.05                       # the returning pulse
        peaks[bc] = dv[i, 0] - 1.0
        # it runs left: dp = -rho c du
        assert np.sign(cv[i, 1]) == -np.sign(peaks[bc])
    assert peaks[rm.BC_SLIPWALL] > 4e-4
    assert peaks[rm.BC_OUTFLOW] / peaks[rm.BC_SLIPWALL] == pytest.approx(-1.0, abs=0.03)
