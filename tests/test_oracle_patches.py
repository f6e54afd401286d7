"""The oracle's total-condition inflow (10), farfield (80) and injection (90) patches,
pinned on the CPU by the physics that fixes each boundary state:

* total-condition inflow (libfloflu/BcondInflowPerf.F90): total temperature and the
  isentropic total pressure of the boundary state, the interior Riemann invariant
  a - (g-1)/2 q.n carried to the boundary, the prescribed flow angles; supersonic
  states from the isentropic Mach-number relations;
* farfield (modflu/RFLU_ModRindStates.F90:248-420): a free stream passes unchanged;
  subsonic faces keep the outgoing acoustic characteristic p + rho a q.n of the
  interior and the incoming one, entropy and tangential velocity of the free stream
  (outflow faces: free-stream pressure, interior entropy and tangential velocity);
* injection (:459-584): injected mass flux, momentum and enthalpy in closed form;
  the wall pressure of a non-burning face from the C- Riemann invariant for a receding
  gas and from Rankine-Hugoniot for the piston shock of an approaching one;
* face fluxes: the central average of the Euler normal fluxes of the interior state
  (with the cell pressure) and the boundary state; injection time-step terms
  (rocflu/RFLU_TimeStepInviscid.F90:214-260) by hand on a one-cell box.

The one branch these leave open is the stagnation fallback of the subsonic inflow
(negative discriminant, which no physical interior state at gamma = 1.4 reaches);
the GPU parity tests cover it no better, so it stays pinned only by reading.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rocflu_amd import mesh as rmesh
from rocflu_amd.solver import Numerics

MM, RG, GP = Numerics().gas_prime()
CPP = GP * RG / (GP - 1.0)        # cp consistent with R' and gamma'
NX = np.array([-1.0, 0.0, 0.0, 2.5])   # outward normal of an x- face, |A| = 2.5


def oracle(bc, opt=0, opt2=0, vals=(0.0,) * 5):
    """Oracle of a 2x2x2 hex box whose patch 0 has type bc (others slip walls)."""
    m = rmesh.box(rmesh.HEX, 2, 2, 2, bc_type=(bc,) + (rmesh.BC_SLIPWALL,) * 5,
                  bc_opt=(opt, 0, 0, 0, 0, 0), stencils=False)
    num = Numerics(patch_vals=[list(vals)] + [[0.0] * 5] * 5,
                   patch_opt2=[opt2, 0, 0, 0, 0, 0])
    return Oracle(m, num)


def cons(rho, u, p):
    u = np.asarray(u, dtype=float)
    return np.array([rho, rho * u[0], rho * u[1], rho * u[2],
                     rho * (p / (rho * (GP - 1.0)) + 0.5 * u @ u)])


def euler_flux(q, p, n):
    """Normal Euler flux of conserved state q with pressure p through face n."""
    rho = q[0]
    u = q[1:4] / rho
    qn = u @ n[:3]
    return np.concatenate([[rho * qn], q[1:4] * qn + p * n[:3], [(q[4] + p) * qn]]) * n[3]


def close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


def direction(bh, bv):
    return np.array([np.cos(bh) * np.cos(bv), np.sin(bh), np.cos(bh) * np.sin(bv)])


# ---------------------------------------------------------------- inflow 10 ----------

TOT = (2.0e5, 320.0, 0.2, -0.1, 1.8)    # ptot, ttot, betah, betav, Mach


@pytest.mark.parametrize("fixed", [0, 1])
@pytest.mark.parametrize("u", [(60.0, 5.0, -3.0), (20.0, -40.0, 10.0), (150.0, 0.0, 0.0),
                               (-30.0, 10.0, 0.0)])
def test_totang_subsonic_state(fixed, u):
    o = oracle(rmesh.BC_INFLOW_TOTANG, opt=1, opt2=fixed, vals=TOT)
    rho, p = 1.2, 1.0e5
    q = cons(rho, u, p)
    r = o.rind_state(0, NX, q, p)
    rr, mr, er, pr = r[0], r[1:4], r[4], r[5]
    tr = pr / (rr * RG)
    ur = mr / rr
    sr = np.linalg.norm(ur)
    ptot, ttot, bh, bv, _ = TOT
    # total temperature and isentropic total pressure
    assert close(tr + sr * sr / (2.0 * CPP), ttot)
    assert close(pr, ptot * (tr / ttot) ** (GP / (GP - 1.0)))
    # prescribed direction and energy
    assert np.allclose(ur, sr * direction(bh, bv), rtol=1e-12, atol=1e-12 * sr)
    assert close(er, pr / (GP - 1.0) + 0.5 * rr * sr * sr)
    # interior Riemann invariant Rm = a - (g-1)/2 q.n, with the boundary normal velocity
    # eta * |u_b| (eta the interior's direction cosine, or -1 for fixed angles).  The
    # reference takes the larger root of the quadratic, so where the interior flows out
    # (q.n > 0) only |a_b - Rm| = (g-1)/2 |eta| |u_b| holds.
    ul = np.asarray(u)
    a02 = GP * RG * ttot
    al = np.sqrt(a02 - 0.5 * (GP - 1.0) * ul @ ul)
    ql = ul @ NX[:3]
    eta = -1.0 if fixed else ql / np.linalg.norm(ul)
    ar = np.sqrt(GP * RG * tr)
    rm = al - 0.5 * (GP - 1.0) * ql
    if ql < 0.0:
        assert close(ar - 0.5 * (GP - 1.0) * eta * sr, rm)
    else:
        assert close(abs(ar - rm), 0.5 * (GP - 1.0) * abs(eta) * sr, 1e-10)


@pytest.mark.parametrize("opt,u", [(0, (60.0, 5.0, 0.0)), (0, (900.0, 0.0, 0.0)),
                                   (2, (900.0, 10.0, 0.0))])
def test_totang_mach_state(opt, u):
    """Supersonic option, and the mixed option where |q.n| >= a: the isentropic state
    of the given Mach number."""
    o = oracle(rmesh.BC_INFLOW_TOTANG, opt=opt, vals=TOT)
    q = cons(1.2, u, 1.0e5)
    r = o.rind_state(0, NX, q, 1.0e5)
    rr, ur, pr = r[0], r[1:4] / r[0], r[5]
    ptot, ttot, bh, bv, mach = TOT
    tr = pr / (rr * RG)
    k = 1.0 + 0.5 * (GP - 1.0) * mach * mach
    assert close(tr * k, ttot)
    assert close(pr * k ** (GP / (GP - 1.0)), ptot)
    assert np.allclose(ur, mach * np.sqrt(GP * RG * tr) * direction(bh, bv), rtol=1e-12)
    assert close(r[4], pr / (GP - 1.0) + 0.5 * rr * ur @ ur)


def test_totang_mixed_subsonic_face_equals_subsonic_option():
    q = cons(1.2, (60.0, 5.0, -3.0), 1.0e5)
    a = oracle(rmesh.BC_INFLOW_TOTANG, opt=2, vals=TOT).rind_state(0, NX, q, 1.0e5)
    b = oracle(rmesh.BC_INFLOW_TOTANG, opt=1, vals=TOT).rind_state(0, NX, q, 1.0e5)
    assert np.array_equal(a, b)


def test_totang_subsonic_option_without_state_is_nan():
    """|q.n| >= a under the subsonic option: the reference leaves the state unset."""
    o = oracle(rmesh.BC_INFLOW_TOTANG, opt=1, vals=TOT)
    r = o.rind_state(0, NX, cons(1.2, (900.0, 0.0, 0.0), 1.0e5), 1.0e5)
    assert np.all(np.isnan(r))


# ---------------------------------------------------------------- farfield 80 --------

def freestream(mach, alpha, beta, p, t):
    rho = p / (RG * t)
    u = mach * np.sqrt(GP * RG * t) * np.array([np.cos(alpha) * np.cos(beta),
                                                 np.sin(alpha) * np.cos(beta), np.sin(beta)])
    return rho, u, p


@pytest.mark.parametrize("mach", [0.4, 2.5])
@pytest.mark.parametrize("nsign", [-1.0, 1.0])
def test_farfield_passes_the_free_stream(mach, nsign):
    vals = (mach, 0.1, -0.05, 1.0e5, 290.0)
    o = oracle(rmesh.BC_FARFIELD, vals=vals)
    rho, u, p = freestream(*vals)
    q = cons(rho, u, p)
    n = np.array([nsign, 0.0, 0.0, 1.0])
    r = o.rind_state(0, n, q, p)
    assert np.allclose(r[:5], q, rtol=1e-12)
    assert close(r[5], p)
    assert np.allclose(o.patch_flux(0, n, q, p), euler_flux(q, p, n), rtol=1e-12, atol=1e-9)


def split(u, n):
    qn = u @ n[:3]
    return qn, u - qn * n[:3]


@pytest.mark.parametrize("u", [(-40.0, 12.0, 3.0), (35.0, -8.0, 0.0)])
def test_farfield_subsonic_characteristics(u):
    vals = (0.5, 0.2, 0.1, 1.0e5, 290.0)
    o = oracle(rmesh.BC_FARFIELD, vals=vals)
    ri, ui, pi = freestream(*vals)
    rl, ul, pl = 1.1, np.asarray(u), 1.2e5
    q = cons(rl, ul, pl)
    r = o.rind_state(0, NX, q, pl)
    rb, ub, pb = r[0], r[1:4] / r[0], r[5]
    al = np.sqrt(GP * pl / rl)
    qb, tb = split(ub, NX)
    ql, tl = split(ul, NX)
    qi, ti = split(ui, NX)
    assert close(r[4], pb / (GP - 1.0) + 0.5 * rb * ub @ ub)
    # the outgoing characteristic (interior side) always holds
    assert close(pb + rl * al * qb, pl + rl * al * ql)
    if ql < 0.0:    # inflow face: the incoming characteristic, entropy, tangential u of the free stream
        assert close(pb - rl * al * qb, pi - rl * al * qi)
        assert close(rb - ri, (pb - pi) / (al * al), 1e-10)
        assert np.allclose(tb, ti, rtol=1e-12, atol=1e-12 * np.linalg.norm(ui))
    else:           # outflow face: free-stream pressure, interior entropy and tangential u
        assert close(pb, pi)
        assert close(rb - rl, (pb - pl) / (al * al), 1e-10)
        assert np.allclose(tb, tl, rtol=1e-12, atol=1e-12 * np.linalg.norm(ul))


def test_farfield_supersonic_inflow_and_outflow():
    vals = (2.0, 0.0, 0.0, 1.0e5, 290.0)
    o = oracle(rmesh.BC_FARFIELD, vals=vals)
    ri, ui, pi = freestream(*vals)
    q_in = cons(1.3, (500.0, 0.0, 0.0), 1.1e5)       # moving into the domain at x-
    r = o.rind_state(0, NX, q_in, 1.1e5)
    assert np.allclose(r[:5], cons(ri, ui, pi), rtol=1e-12) and close(r[5], pi)
    q_out = cons(1.3, (-500.0, 20.0, 0.0), 1.1e5)    # leaving through x-
    r = o.rind_state(0, NX, q_out, 1.1e5)
    assert np.array_equal(r[:5], q_out)
    assert close(r[5], 1.1e5)      # pressure formed from rho E


@pytest.mark.parametrize("bc,opt,vals,u", [
    (rmesh.BC_FARFIELD, 0, (0.5, 0.2, 0.1, 1.0e5, 290.0), (-40.0, 12.0, 3.0)),
    (rmesh.BC_FARFIELD, 0, (0.5, 0.2, 0.1, 1.0e5, 290.0), (35.0, -8.0, 0.0)),
    (rmesh.BC_INFLOW_TOTANG, 1, TOT, (60.0, 5.0, -3.0)),
    (rmesh.BC_INFLOW_TOTANG, 0, TOT, (60.0, 5.0, -3.0)),
])
def test_central_patch_flux(bc, opt, vals, u):
    """Face flux = average of the Euler fluxes of the interior state, carrying the
    cell pressure, and of the boundary state."""
    o = oracle(bc, opt=opt, vals=vals)
    q = cons(1.1, u, 1.2e5)
    pl = 1.2e5 * (1.0 + 1e-3)        # a cell pressure apart from rho E's, to tell them apart
    r = o.rind_state(0, NX, q, pl)
    want = 0.5 * (euler_flux(q, pl, NX) + euler_flux(r[:5], r[5], NX))
    got = o.patch_flux(0, NX, q, pl)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


# ---------------------------------------------------------------- injection 90 -------

@pytest.mark.parametrize("minj,tinj", [(2.5, 2800.0), (0.01, 500.0)])
def test_injection_burning_face(minj, tinj):
    o = oracle(rmesh.BC_INJECTION, vals=(minj, tinj, 0.0, 0.0, 0.0))
    pl = 3.0e5
    q = cons(1.0, (5.0, 1.0, 0.0), 1.0e5)
    r = o.rind_state(0, NX, q, pl)
    rho = pl / (RG * tinj)
    assert close(r[0], rho)
    assert np.allclose(r[1:4], -(minj / rho) * NX[:3], rtol=1e-12, atol=0.0)
    assert close(r[4], Numerics().cp * tinj + 0.5 * (minj / rho) ** 2)
    f = o.patch_flux(0, NX, q, pl)
    nm = NX[3]
    assert f[0] == -minj * nm
    assert np.allclose(f[1:4], (minj * minj / rho + pl) * NX[:3] * nm, rtol=1e-12, atol=0.0)
    assert close(f[4], -minj * (Numerics().cp * tinj + 0.5 * (minj / rho) ** 2) * nm)


def hugoniot_velocity_jump(p1, rho1, p2):
    """Velocity change across a shock from state (p1, rho1) to pressure p2."""
    g = GP
    rho2 = rho1 * ((g + 1.0) * p2 + (g - 1.0) * p1) / ((g - 1.0) * p2 + (g + 1.0) * p1)
    j = np.sqrt((p2 - p1) / (1.0 / rho1 - 1.0 / rho2))
    return j * (1.0 / rho1 - 1.0 / rho2)


@pytest.mark.parametrize("un", [-120.0, -10.0, 0.0, 10.0, 120.0, 900.0])
def test_injection_wall_pressure(un):
    """Non-burning face: the gas's normal velocity un = u.n (n outward) is stopped at
    the wall."""
    o = oracle(rmesh.BC_INJECTION, vals=(0.0, 0.0, 0.0, 0.0, 0.0))
    rho, p = 1.2, 1.0e5
    u = un * NX[:3] + np.array([0.0, 30.0, -5.0])     # plus tangential velocity
    q = cons(rho, u, p)
    pw = o.rind_state(0, NX, q, p)[5]
    a = np.sqrt(GP * p / rho)
    if un < 0.0:     # receding gas: C- invariant a + (g-1)/2 u.n, isentropic
        aw = a + 0.5 * (GP - 1.0) * un
        assert close(pw, p * (aw / a) ** (2.0 * GP / (GP - 1.0)), 1e-11)
    elif un == 0.0:
        assert pw == p
    else:            # approaching gas: the piston-shock pressure brings it to rest
        assert pw > p
        assert close(hugoniot_velocity_jump(p, rho, pw), un, 1e-10)
    f = o.patch_flux(0, NX, q, p)
    assert f[0] == 0.0 and f[4] == 0.0
    assert np.allclose(f[1:4], pw * NX[:3] * NX[3], rtol=1e-15, atol=0.0)


def test_injection_time_step_terms():
    """One hexahedron: x- burning injection, x+ non-burning injection, slip walls;
    dt = vol / sum of face terms, by hand."""
    m = rmesh.box(rmesh.HEX, 1, 1, 1, bc_type=(rmesh.BC_INJECTION, rmesh.BC_INJECTION) +
                  (rmesh.BC_SLIPWALL,) * 4, stencils=False)
    minj, tinj = 2.0, 2500.0
    num = Numerics(patch_vals=[[minj, tinj, 0, 0, 0], [0.0] * 5] + [[0.0] * 5] * 4)
    o = Oracle(m, num)
    rho, u, p = 1.1, np.array([30.0, -20.0, 7.0]), 1.5e5
    cv = cons(rho, u, p)[None, :]
    dv = o.dv(cv)
    a = dv[0, 2]
    srad = 0.0
    for b in range(6):
        n = m.bfn[b]
        if b == 0:
            rho_inj = dv[0, 0] / (RG * tinj)
            srad += (abs(-minj / rho_inj) + np.sqrt(GP * RG * tinj)) * n[3]
        elif b == 1:
            srad += a * n[3]
        else:
            srad += (abs(u @ n[:3]) + a) * n[3]
    assert close(o.local_dt(cv, dv), m.vol[0] / srad, 1e-14)


def test_planted_mistakes_fail_the_pins():
    """The invariants above reject plausible slips: the wrong sign or a dropped half in
    the inflow Riemann invariant (fixed angles, interior flow with a tangential part),
    and a missing factor in the piston-shock pressure."""
    ttot = TOT[1]
    ul = np.array([60.0, 5.0, -3.0])
    a02 = GP * RG * ttot
    al = np.sqrt(a02 - 0.5 * (GP - 1.0) * ul @ ul)
    ql = ul @ NX[:3]
    eta = -1.0
    k = 0.5 * (GP - 1.0) * eta * eta
    for name, rm in (("right", al - 0.5 * (GP - 1.0) * ql), ("sign", al + 0.5 * (GP - 1.0) * ql),
                     ("half", al - (GP - 1.0) * ql)):
        disc = k * (a02 / (rm * rm) * (1.0 + k) - 1.0)
        ar = rm / (1.0 + k) * (1.0 + np.sqrt(disc))
        sr = np.sqrt(2.0 / (GP - 1.0) * (a02 - ar * ar))
        holds = close(ar - 0.5 * (GP - 1.0) * eta * sr, al - 0.5 * (GP - 1.0) * ql)
        assert holds == (name == "right"), name
    rho, p, un = 1.2, 1.0e5, 120.0
    a = np.sqrt(GP * p / rho)
    t = (GP + 1.0) / 4.0
    good = p + t * rho * un * (un + np.sqrt(a * a + t * t * un * un) / t)
    bad = p + t * rho * un * (un + np.sqrt(a * a + t * t * un * un))
    assert close(hugoniot_velocity_jump(p, rho, good), un, 1e-10)
    assert not close(hugoniot_velocity_jump(p, rho, bad), un, 1e-3)
