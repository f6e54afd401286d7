"""Pins of the CPU oracle by what the mathematics fixes, independently of how the
oracle writes its formulas: consistency and upwinding of the flux functions, the
stationary contact of HLLC, exactness of the WLSQ gradient on linear fields and its
equality with a weighted least-squares solve, the RK stability polynomials, the EOS,
the time-step formula on a uniform box, the after-update floor, and Sod's problem
against the exact Riemann solution."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rocflu_amd import cases, mesh as rm
from rocflu_amd.solver import Numerics


def two_cells():
    return rm.box(rm.HEX, 2, 1, 1, lengths=(2.0, 1.0, 1.0), jitter=0.0,
                  bc_type=(rm.BC_SLIPWALL,) * 6, stencils=False)


def phys_flux_x(rho, u, v, w, p, gp):
    E = p / (gp - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.array([rho * u, rho * u * u + p, rho * u * v, rho * u * w, u * (E + p)])


def interior_flux(discr, left, right, gamma=1.4, cp=1004.5):
    """Flux through the single interior face (normal +x, |A| = 1) of a 2-cell box:
    the residual of cell 0 minus its slip-wall pressure terms."""
    m = two_cells()
    num = Numerics(discr=discr, order=1, gamma=gamma, cp=cp)
    gp = num.gas_prime()[2]
    cv = np.array([cases.prim_to_cons(*s, gp) for s in (left, right)])
    o = Oracle(m, num)
    rhs = o.residual(cv)
    p0 = o.dv(cv)[0, 0]
    # walls of cell 0: x- (normal -x) and the y/z pairs, which cancel
    return rhs[0] + np.array([0.0, p0, 0.0, 0.0, 0.0]), gp


@pytest.mark.parametrize("discr", [1, 3, 4])
def test_flux_consistency(discr):
    """F(U, U) equals the Euler flux f(U).n |A| for Roe, HLLC and AUSM+."""
    for s in [(1.0, 0.3, -0.2, 0.1, 1.0), (0.5, -1.7, 0.4, 0.0, 2.5), (2.0, 0.0, 0.0, 0.0, 0.7)]:
        f, gp = interior_flux(discr, s, s)
        assert np.allclose(f, phys_flux_x(*s, gp), rtol=1e-13, atol=1e-13), (discr, s)


@pytest.mark.parametrize("discr", [1, 3, 4])
def test_supersonic_upwinding(discr):
    """Both states supersonic to the right: the flux is the left state's Euler flux
    (Roe's dissipation cancels exactly, HLLC takes the left branch, AUSM+ the upwind
    Mach split)."""
    left = (1.0, 3.0, 0.1, 0.0, 1.0)
    right = (0.8, 2.6, -0.2, 0.1, 0.7)
    f, gp = interior_flux(discr, left, right)
    assert np.allclose(f, phys_flux_x(*left, gp), rtol=1e-12, atol=1e-12), discr


def test_hllc_resolves_stationary_contact():
    """A contact at rest (equal p, u = 0, different rho) gives no mass or energy flux."""
    f, _ = interior_flux(3, (1.0, 0.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.0, 1.0))
    assert np.allclose(f, [0.0, 1.0, 0.0, 0.0, 0.0], atol=1e-14)


def test_roe_is_dissipative_across_contact():
    """Roe smears the same contact (nonzero mass flux from |A| dU): the dissipative
    term is present and has the upwind sign (mass moves toward the lighter side is
    not physical here; the flux must be nonzero and antisymmetric under L<->R)."""
    a, _ = interior_flux(1, (1.0, 0.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.0, 1.0))
    b, _ = interior_flux(1, (0.125, 0.0, 0.0, 0.0, 1.0), (1.0, 0.0, 0.0, 0.0, 1.0))
    assert abs(a[0]) > 1e-3
    assert np.isclose(a[0], -b[0], rtol=1e-12)


def linear_field(mesh, rng):
    c = mesh.cofg
    coef = rng.normal(size=(5, 4))
    return np.stack([coef[k, 0] + c @ coef[k, 1:] for k in range(5)], axis=1), coef


@pytest.mark.parametrize("kind", [rm.HEX, rm.TET6])
def test_wlsq_gradient_exact_on_linear_fields(kind):
    """The reference's own check (RFLU_ModDifferentiationCells.F90:575-584): linear
    data give the exact gradient in every cell, boundary cells included."""
    m = rm.box(kind, 5, 4, 3, lengths=(5.0, 4.0, 3.0), jitter=0.1, seed=3)
    var, coef = linear_field(m, np.random.default_rng(1))
    num = Numerics(order=2, reconst=0)
    g = Oracle(m, num).gradients(var)
    assert np.allclose(g, np.broadcast_to(coef[None, :, 1:], g.shape), rtol=1e-10, atol=1e-10)


def test_wlsq_gradient_equals_weighted_lstsq():
    """On nonlinear data the gradient is the least-squares fit with rows weighted by
    1/|d| (the R-factor form of the normal equations)."""
    m = rm.box(rm.TET6, 4, 4, 4, lengths=(4.0, 4.0, 4.0), jitter=0.08, seed=9)
    rng = np.random.default_rng(2)
    var = rng.normal(size=(m.n_cells, 5))
    g = Oracle(m, Numerics(order=2)).gradients(var)
    for i in rng.choice(m.n_cells, 25, replace=False):
        mem = m.c2cs_idx[m.c2cs_ptr[i]:m.c2cs_ptr[i + 1]]
        d = m.cofg[mem] - m.cofg[i]
        w = 1.0 / np.linalg.norm(d, axis=1)
        for k in range(5):
            sol = np.linalg.lstsq(d * w[:, None], (var[mem, k] - var[i, k]) * w, rcond=None)[0]
            assert np.allclose(g[i, k], sol, rtol=1e-10, atol=1e-12)


def test_weno_and_clip_properties():
    """WENO-XYZ keeps a uniform gradient field and zeroes a positive variable's gradient
    whose face extrapolation would be non-positive."""
    m = rm.box(rm.HEX, 4, 4, 4, lengths=(4.0, 4.0, 4.0), jitter=0.05, seed=4)
    o = Oracle(m, Numerics(order=2, reconst=2))
    g = np.tile(np.arange(1.0, 16.0), (m.n_cells, 1))
    assert np.allclose(o.weno_xyz(g), g, rtol=1e-14)
    var = np.ones((m.n_cells, 5))
    var[:, 0] = 0.01
    grad = np.zeros((m.n_cells, 15))
    grad[:, 0] = 1.0       # d(rho)/dx = 1 drives the x- face value negative
    grad[:, 3] = 1.0       # u is not positive-definite: kept
    out = o.limit_simple(var, grad)
    assert np.all(out[:, 0:3] == 0.0) and np.all(out[:, 3] == 1.0)


@pytest.mark.parametrize("rk,poly", [
    (2, [1, -1, 1 / 2, -1 / 6]),
    (1, [1, -1, 1 / 2, -1 / 6, 1 / 24]),
    (4, [1, -1, 1 / 2, -1 / 6]),
    (3, [1, -1]),
])
def test_rk_stability_polynomial(rk, poly):
    """One step on du/dt = -lambda u reproduces the scheme's polynomial in z = lambda dt
    (Wray and SSP-RK3: third order; classical RK4: fourth; Euler: first)."""
    m = two_cells()
    o = Oracle(m, Numerics(rk=rk))
    lam, dt = 0.7, 0.3
    u0 = np.array([[1.0, 2.0, -0.5, 0.25, 3.0], [0.5, 1.0, 2.0, -1.0, 0.1]])
    cv = u0.copy()
    cv_old = u0.copy()
    rhs_sum = np.zeros_like(u0)
    nst = {1: 4, 2: 3, 3: 1, 4: 3}[rk]
    for s in range(1, nst + 1):
        if rk in (2, 3):
            cv_old = cv.copy()
        if rk == 4:
            rhs_sum = cv.copy()
        rhs = lam * cv * m.vol[:, None]
        cv = o.rk_update(s, dt, cv.copy(), cv_old, rhs, rhs_sum)
    z = lam * dt
    expect = sum(c * z ** k for k, c in enumerate(poly))
    assert np.allclose(cv, u0 * expect, rtol=1e-14)


def test_dependent_vars_closed_form():
    m = two_cells()
    num = Numerics(gamma=1.35, cp=1100.0)
    mm, rg, gp = num.gas_prime()
    cv = np.array([[1.2, 0.3, -0.6, 0.9, 250000.0], [0.4, 0.0, 0.0, 0.0, 90000.0]])
    dv = Oracle(m, num).dv(cv)
    u = cv[:, 1:4] / cv[:, :1]
    p = (gp - 1.0) * (cv[:, 4] - 0.5 * cv[:, 0] * (u * u).sum(1))
    t = p / (cv[:, 0] * rg)
    assert np.allclose(dv[:, 0], p, rtol=1e-14)
    assert np.allclose(dv[:, 1], t, rtol=1e-14)
    assert np.allclose(dv[:, 2], np.sqrt(gp * rg * t), rtol=1e-14)
    # gamma' is gamma up to the M <-> R round trip
    assert abs(gp - 1.35) < 1e-13


def test_time_step_uniform_box():
    """Unit cubes, uniform flow u along x: dt = vol / (2(|u| + c) + 4c), then the CFL,
    imposed-first-step and DTMINLIMIT rules."""
    m = rm.box(rm.HEX, 3, 3, 3, lengths=(3.0, 3.0, 3.0), jitter=0.0,
               bc_type=(rm.BC_SLIPWALL,) * 6, stencils=False)
    num = Numerics(cfl=0.5, dt_imposed=1e-3, dt_min_limit=0.0)
    gp = num.gas_prime()[2]
    u, rho, p = 0.4, 1.3, 0.9
    cv = np.tile(cases.prim_to_cons(rho, u, 0.0, 0.0, p, gp), (m.n_cells, 1))
    o = Oracle(m, num)
    c = np.sqrt(gp * p / rho)
    dtl = o.local_dt(cv)
    assert np.isclose(dtl, 1.0 / (2 * (u + c) + 4 * c), rtol=1e-13)
    assert np.isclose(o.min_timestep(dtl, 1.0), 0.5 * dtl, rtol=1e-14)
    assert np.isclose(o.min_timestep(dtl, 0.0), 1e-3, rtol=1e-15)
    num2 = Numerics(cfl=0.5, dt_imposed=1e-3, dt_min_limit=1.0)
    assert Oracle(m, num2).min_timestep(dtl, 1.0) == 1.0


def test_after_update_floor_and_errors():
    m = two_cells()
    num = Numerics()
    o = Oracle(m, num)
    gp = num.gas_prime()[2]
    cv = np.array([cases.prim_to_cons(1.0, 0.5, 0.0, 0.0, 1.0, gp),
                   cases.prim_to_cons(1.0, 0.5, 0.0, 0.0, 1.0, gp)])
    cv[1, 4] = 0.5 * 1.0 * 0.25 - 0.01      # internal energy < 0: p < 0
    dv = o.dv(cv)
    rc, cv2, dv2, st = o.after_update(cv, dv)
    assert rc == 0 and st["n_pos_fix"] == 1
    assert np.isclose(cv2[1, 4], 1.0 * (1.0e4 + 0.5 * 0.25))
    assert dv2[1, 0] > 0
    cv[0, 2] = np.nan
    rc, *_ = o.after_update(cv, o.dv(cv))
    assert rc == 1  # ERR_INVALID_VALUE


def exact_sod(x, t, x0, L=(1.0, 0.0, 1.0), R=(0.125, 0.0, 0.1), g=1.4):
    """Exact Riemann solution (Toro, ch. 4) for two gases at rest."""
    rl, ul, pl = L
    rr, ur, pr = R
    al, ar = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)

    def fk(p, rk, pk, ak):
        if p > pk:
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ak / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-8, 10.0
    for _ in range(200):
        pm = 0.5 * (lo + hi)
        if fk(pm, rl, pl, al) + fk(pm, rr, pr, ar) + ur - ul > 0:
            hi = pm
        else:
            lo = pm
    um = 0.5 * (ul + ur) + 0.5 * (fk(pm, rr, pr, ar) - fk(pm, rl, pl, al))
    rml = rl * (pm / pl) ** (1 / g)
    rmr = rr * ((pm / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * pm / pr + 1))
    sh = ur + ar * np.sqrt((g + 1) / (2 * g) * pm / pr + (g - 1) / (2 * g))
    aml = al * (pm / pl) ** ((g - 1) / (2 * g))
    s = (x - x0) / t
    rho = np.where(s < ul - al, rl, 0.0)
    fan = (s >= ul - al) & (s < um - aml)
    rho = np.where(fan, rl * (2 / (g + 1) + (g - 1) / ((g + 1) * al) * (ul - s)) ** (2 / (g - 1)), rho)
    rho = np.where((s >= um - aml) & (s < um), rml, rho)
    rho = np.where((s >= um) & (s < sh), rmr, rho)
    rho = np.where(s >= sh, rr, rho)
    return rho


def test_sod_converges_to_exact_solution():
    """Density L1 error against the exact solution is small and shrinks under mesh
    refinement for every flux (first order: roughly halves per doubling)."""
    for discr in (1, 3, 4):
        errs = []
        for n in (100, 200, 400):
            m = rm.box(rm.HEX, n, 1, 1, lengths=(1.0, 1.0 / n, 1.0 / n), jitter=0.0,
                       bc_type=(rm.BC_OUTFLOW, rm.BC_OUTFLOW) + (rm.BC_SLIPWALL,) * 4,
                       stencils=False)
            num = Numerics(discr=discr, order=1, rk=2, cfl=0.5, dt_imposed=1e-4)
            cv0 = cases.sod(m, num, 0.5)
            cv, dv, st = Oracle(m, num).run_until(cv0, 0.15)
            assert st["rc"] == 0
            ex = exact_sod(m.cofg[:, 0], st["time"], 0.5, g=num.gas_prime()[2])
            errs.append(np.abs(cv[:, 0] - ex).mean())
        assert errs[1] < 0.02, (discr, errs)
        assert errs[2] < 0.75 * errs[1] < 0.75 * 0.75 * errs[0], (discr, errs)


def test_closed_box_conserves_mass_and_energy():
    m = rm.box(rm.TET6, 4, 4, 4, lengths=(4.0, 4.0, 4.0), jitter=0.08, seed=8,
               bc_type=(rm.BC_SLIPWALL,) * 6)
    for order, reconst in ((1, 0), (2, 11), (2, 2)):
        num = Numerics(discr=3, order=order, reconst=reconst, rk=2, cfl=0.5, dt_imposed=1e-3)
        cv0 = cases.planar_shock(m, num, 1.5, mach=2.0, noise=0.01, seed=1)
        cv, _, st = Oracle(m, num).run(cv0, 20)
        assert st["rc"] == 0
        for k in (0, 4):
            assert abs((cv[:, k] * m.vol).sum() / (cv0[:, k] * m.vol).sum() - 1) < 1e-13


def inflow_box(order=1, stencils=False):
    """4x3x3 jittered hexes: inflow (velocity + temperature) at x-, supersonic outflow
    at x+, slip walls elsewhere.  Boundary nodes stay on the box planes."""
    return rm.box(rm.HEX, 4, 3, 3, lengths=(4.0, 3.0, 3.0), jitter=0.1, seed=3,
                  bc_type=(rm.BC_INFLOW_VELTEMP, rm.BC_OUTFLOW) + (rm.BC_SLIPWALL,) * 4,
                  bc_opt=(0,) * 6, stencils=stencils)


def normal_flux(rho, u, p, gp, n):
    """Physical Euler flux F(U).n per unit area for velocity vector u."""
    q = float(np.dot(u, n))
    E = p / (gp - 1.0) + 0.5 * rho * float(np.dot(u, u))
    return np.array([rho * q, rho * u[0] * q + p * n[0], rho * u[1] * q + p * n[1],
                     rho * u[2] * q + p * n[2], (E + p) * q])


@pytest.mark.parametrize("order,opt", [(1, 1), (1, 0), (2, 1), (2, 0)])
def test_inflow_preserves_free_stream(order, opt):
    """A uniform stream whose temperature, velocity (and pressure) the inflow patch
    prescribes has zero residual: the boundary state equals the interior state.  (At
    order 2 every patch runs at order 2: a first-order outflow patch in a second-order
    run reads p in the energy slot, as the reference does, which no uniform stream
    survives.)"""
    m = inflow_box(order, stencils=order == 2)
    rho, u, p = 1.2, 0.3, 1.0
    base = Numerics(discr=3, order=order, reconst=0)
    mm, rg, gp = base.gas_prime()
    T = p / (rho * rg)
    po = [2] * 6 if order == 2 else None
    num = Numerics(discr=3, order=order, reconst=0, patch_order=po,
                   patch_vals=[[u, 0.0, 0.0, T, p]] + [[0.0] * 5] * 5)
    opts = np.array(m.a["patch_opt"]); opts[0] = opt; m.a["patch_opt"] = opts
    cv = np.tile(cases.prim_to_cons(rho, u, 0.0, 0.0, p, gp), (m.n_cells, 1))
    r = Oracle(m, num).residual(cv)
    assert np.abs(r).max() <= 1e-13 * p, np.abs(r).max()


@pytest.mark.parametrize("opt", [1, 0])
def test_inflow_boundary_state_closed_form(opt):
    """Uniform interior state, different inflow values: the residual of each cell on
    the inflow patch is half the jump of the physical flux between the boundary state
    (subsonic: interior density, p = rho R T_in; otherwise p_in and rho = p_in/(R
    T_in)) and the interior state, times the face area; elsewhere it vanishes."""
    m = inflow_box()
    rho, u0, p = 1.2, 0.3, 1.0
    num0 = Numerics(discr=3)
    mm, rg, gp = num0.gas_prime()
    T_in, u_in, p_in = 1.3 * p / (rho * rg), np.array([0.5, 0.1, -0.05]), 2.0
    num = Numerics(discr=3, patch_vals=[[*u_in, T_in, p_in]] + [[0.0] * 5] * 5)
    opts = np.array(m.a["patch_opt"]); opts[0] = opt; m.a["patch_opt"] = opts
    cv = np.tile(cases.prim_to_cons(rho, u0, 0.0, 0.0, p, gp), (m.n_cells, 1))
    r = Oracle(m, num).residual(cv)
    if opt == 1:
        rb, pb = rho, rho * rg * T_in
    else:
        rb, pb = p_in / (rg * T_in), p_in
    expect = np.zeros_like(r)
    bp = m.a["bpatch_ptr"]
    for b in range(bp[0], bp[1]):
        c = m.a["bf2c"][b]
        n, area = m.a["bfn"][b, :3], m.a["bfn"][b, 3]
        expect[c] += 0.5 * (normal_flux(rb, u_in, pb, gp, n) -
                            normal_flux(rho, np.array([u0, 0.0, 0.0]), p, gp, n)) * area
    assert np.abs(r - expect).max() <= 1e-13 * np.abs(expect).max(), np.abs(r - expect).max()


@pytest.mark.parametrize("kind", [rm.HEX, rm.TET6])
def test_constrained_gradient_exact_on_admissible_linear_fields(kind):
    """CRECONSTC 1: a cell's velocity gradient is refitted with the slip-wall rows
    (u.n = 0 on walls normal to the axis).  For u = a x (zero on the x = 0 wall) the
    fit of every cell in the lower half of the box is exact, as it is for v = b y and
    w = c z; density and pressure keep their least-squares gradients."""
    n = 6
    m = rm.box(kind, n, n, n, lengths=(float(n),) * 3, jitter=0.1, seed=2,
               bc_type=(rm.BC_SLIPWALL,) * 6)
    num = Numerics(discr=3, order=2, reconst=0, creconst=1)
    x = m.cofg
    var = np.empty((m.n_cells, 5))
    var[:, 0] = 1.0 + 0.1 * x[:, 0] - 0.2 * x[:, 1] + 0.05 * x[:, 2]
    var[:, 1] = 0.3 * x[:, 0]
    var[:, 2] = -0.2 * x[:, 1]
    var[:, 3] = 0.7 * x[:, 2]
    var[:, 4] = 2.0 - 0.1 * x[:, 2]
    o = Oracle(m, num)
    g = o.gradients(var)
    gc = o.gradients_constr(var, g)
    nb = np.diff(m.a["c2cs_bptr"])
    lower = (x < 0.5 * n).all(axis=1)
    assert (nb[lower] > 0).sum() > 10
    exact = np.array([[0.1, -0.2, 0.05], [0.3, 0, 0], [0, -0.2, 0], [0, 0, 0.7], [0, 0, -0.1]])
    err = np.abs(gc[lower] - exact[None]).max()
    assert err <= 1e-12, err
    # rho and p untouched; velocity rows changed somewhere (the walls act)
    assert np.array_equal(gc[:, [0, 4]], g[:, [0, 4]])
    u = var.copy()
    u[:, 1] = 1.0                                   # violates u = 0 at x walls
    g1 = o.gradients(u)
    g2 = o.gradients_constr(u, g1)
    near = nb > 0
    assert np.abs(g1[near, 1]).max() <= 1e-12 and np.abs(g2[near, 1, 0]).max() > 1e-3


def test_positivity_check_floors_every_cell_before_stopping():
    """RFLU_CheckPositivity (libflu/RFLU_CheckPositivity.F90:147-253) floors the energy of
    every cell with p <= 0 and counts the cells still negative; the error stop comes
    after the loop, so a cell after the first failing one is still floored."""
    m = rm.box(rm.HEX, 3, 1, 1, lengths=(3.0, 1.0, 1.0), jitter=0.0,
               bc_type=(rm.BC_SLIPWALL,) * 6, stencils=False)
    num = Numerics()
    o = Oracle(m, num)
    gp = num.gas_prime()[2]
    cv = np.array([cases.prim_to_cons(1.0, 0.5, 0.0, 0.0, 1.0, gp) for _ in range(3)])
    cv[0, 0] = -0.2                                  # negative density: the error
    cv[2, 4] = 0.5 * 1.0 * 0.25 - 0.01               # p < 0 after it: floored
    rc, cv2, dv2, st = o.after_update(cv, o.dv(cv))
    assert rc == 2                                   # ERR_NEGATIVE_POSDEF
    assert st["n_pos_fix"] == 1                      # cell 2 only (cell 0 has p > 0)
    assert np.isclose(cv2[2, 4], 1.0 * (1.0e4 + 0.5 * 0.25)) and dv2[2, 0] > 0
